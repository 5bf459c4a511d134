"""Post-filters on the GPU (decoder-exact deblocking and CDEF after the closed loop; the reference's GpConfig lfMode /
lfMul / lfLevel / cdef knobs, encoder/src/pipeline.h:60-66 and pipeline.cpp:960-986): the GPU's filtered reconstruction
equals the reference backend's post-filtered reconstruction of the GPU's own record (oracle/_ref: the reference's
av1core post-filters, encoder.cpp:1038-1067) and dav1d's decode of this build's AVIF, on 8-bit 4:2:0, the odd-sized gate
picture and 10-bit 4:4:4, alone and in a batch of frames at different qindex (one launch per pass for the batch);
the post-filters leave every decision unchanged; the frame's levels follow the reference's LPF_PICK_FROM_Q fit."""
import os
import re

import numpy as np
import pytest

import av1gpu_amd as A
import refcheck
from test_gate_image import gate_planes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ac_q(q, bd):
    """Ac_Qlookup of the spec (the table this build ships, identical to the reference's headers by test_tables)"""
    txt = open(os.path.join(ROOT, "av1gpu_amd", "csrc", "av1_spec_tables.h")).read()
    body = txt[txt.index("Ac_Qlookup[3][256]"):]
    body = body[body.index("{") + 1:body.index("};")]
    vals = [int(v) for v in re.findall(r"\d+", body)]
    assert len(vals) == 3 * 256
    return vals[((bd - 8) >> 1) * 256 + q]


def lf_level(q, bd, mode, mul=1.0, level=0):
    """libaom's LPF_PICK_FROM_Q key-frame fit (pipeline.cpp:960-976), restated"""
    if mode == 2:
        return level
    acq = ac_q(q, bd)
    v = (((acq * 20723 + 4060632 + (1 << 19)) >> 20) - 4) if bd == 10 else ((acq * 17563 - 421574 + (1 << 17)) >> 18)
    return int(max(0, min(63, np.rint(np.float32(v) * np.float32(mul)))))


def _check(ref, g, planes, W, H, q, ssx, bd, plain):
    for k in ("blocks", "block_txb", "txb", "coeffs"):
        assert np.array_equal(getattr(g, k), getattr(plain, k)), "the post-filters changed the record: " + k
    dec, _, _ = ref.encode(dict(blocks=g.blocks, block_txb=g.block_txb, txb=g.txb, coeffs=g.coeffs), W, H, q, ssx=ssx, bd=bd,
                           post=g.fp)
    for p in range(3):
        assert np.array_equal(g.recon[p][:dec[p].shape[0], :dec[p].shape[1]].astype(np.uint16), dec[p].astype(np.uint16)), \
            "reference post-filtered plane %d" % p
    changed = any(not np.array_equal(g.recon[p], plain.recon[p]) for p in range(3))
    if refcheck.dav1d_available():
        d = refcheck.dav1d_decode_yuv(A.write_bitstream(g))
        for p in range(3):
            assert np.array_equal(d[p].astype(np.uint16), dec[p].astype(np.uint16)), "dav1d plane %d" % p
    return changed


CONFIGS = [dict(lfMode=1, cdef=1, cdefDampingMinus3=0, cdefYPri=4, cdefYSec=2, cdefUVPri=2, cdefUVSec=1),
           dict(lfMode=2, lfLevel=63, cdef=1, cdefDampingMinus3=3, cdefYPri=15, cdefYSec=4, cdefUVPri=15, cdefUVSec=4),
           dict(lfMode=1, lfMul=1.5, cdef=1, cdefDampingMinus3=2, cdefYPri=7, cdefYSec=1, cdefUVPri=3, cdefUVSec=0),
           dict(lfMode=1), dict(cdef=1, cdefDampingMinus3=1, cdefYPri=9, cdefYSec=2, cdefUVPri=5, cdefUVSec=1)]
CASES = [(200, 136, 3, 120, 8, 1), (211, 149, 0, 60, 8, 1), (136, 72, 50, 160, 10, 0), (264, 200, 6, 200, 8, 1)]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "-".join("%s%s" % (k, v) for k, v in c.items() if k in ("lfMode", "cdefYPri")))
def test_gpu_postfilters_equal_reference_and_dav1d(ref, cfg):
    changed = 0
    for W, H, seed, q, bd, ssx in CASES:
        print("post-filters %s on %dx%d q%d %d-bit %s" % (cfg, W, H, q, bd, "4:2:0" if ssx else "4:4:4"), flush=True)
        planes = gate_planes() if seed == 0 else A.synth_image(seed, W, H, ssx=ssx, bd=bd)
        f = A.DeviceFrame(planes, W, H, bd=bd, qindex=q, yuv444=not ssx)
        e = A.GpEncoder().init(A.GpConfig(qindex=q, **cfg))
        e0 = A.GpEncoder().init(A.GpConfig(qindex=q))
        g, t = e.encode([f])
        plain = e0.encode([f])[0][0]
        e.close()
        e0.close()
        f.close()
        g = g[0]
        lv = lf_level(q, bd, cfg.get("lfMode", 0), cfg.get("lfMul", 1.0), cfg.get("lfLevel", 0)) if cfg.get("lfMode") else 0
        assert list(g.fp.loop_filter_level) == [lv] * 4 and g.fp.enable_cdef == cfg.get("cdef", 0)
        if cfg.get("cdef"):
            assert (g.fp.cdef_damping_minus_3, g.fp.cdef_y_pri[0], g.fp.cdef_y_sec[0], g.fp.cdef_uv_pri[0], g.fp.cdef_uv_sec[0]) == \
                (cfg["cdefDampingMinus3"], cfg["cdefYPri"], cfg["cdefYSec"], cfg["cdefUVPri"], cfg["cdefUVSec"])
        assert t.postfilter > 0
        changed += _check(ref, g, planes, W, H, q, ssx, bd, plain)
    assert changed >= 2, "the post-filters changed almost nothing"


def test_gpu_postfilters_in_a_batch(ref):
    specs = [(1000, 712, 7101, 80), (328, 200, 7102, 160), (4160, 192, 7103, 120), (72, 40, 7104, 255)]
    print("post-filter batch", specs, flush=True)
    planes = [A.synth_image(s, W, H) for W, H, s, q in specs]
    frames = [A.DeviceFrame(p, W, H, qindex=q) for p, (W, H, s, q) in zip(planes, specs)]
    cfg = CONFIGS[0]
    e = A.GpEncoder().init(A.GpConfig(qindex=120, **cfg))
    together, _ = e.encode(frames)
    e0 = A.GpEncoder().init(A.GpConfig(qindex=120))
    plain, _ = e0.encode(frames)
    for k, (p, (W, H, s, q)) in enumerate(zip(planes, specs)):
        alone = e.encode([frames[k]])[0][0]
        for p_ in range(3):
            assert np.array_equal(alone.recon[p_], together[k].recon[p_]), "batch vs alone, plane %d" % p_
        assert list(together[k].fp.loop_filter_level) == [lf_level(q, 8, 1)] * 4
        _check(ref, together[k], p, W, H, q, 1, 8, plain[k])
    e.close()
    e0.close()
    for f in frames:
        f.close()


def test_postfilter_options_are_checked():
    for bad in (dict(lfMode=3), dict(lfMode=2, lfLevel=64), dict(cdef=1, cdefYSec=3), dict(cdef=1, cdefYPri=16),
                dict(cdef=1, cdefDampingMinus3=4)):
        with pytest.raises(RuntimeError, match="post-filter"):
            A.GpEncoder().init(A.GpConfig(qindex=120, **bad))
