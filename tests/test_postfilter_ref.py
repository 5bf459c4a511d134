"""The post-filters' checker and this build's writer on the CPU.  The reference's decoder-exact post-filters
(av1core/include/av1g/postfilter.h, applied by its CPU backend in Encoder::postfilter, av1core/src/encoder.cpp:1038-1067,
built in oracle/_ref) are pinned by dav1d (Pillow's libavif): the reference writer's AVIF of an oracle record with loop
filter and CDEF parameters decodes to the reference backend's post-filtered reconstruction.  This build's writer codes the
same parameters: its AVIF decodes in dav1d to that reconstruction too, and the reference parser reads the parameters
back.  With every level and strength 0 nothing changes; with filters on the reconstruction does change."""
import ctypes as C

import numpy as np
import pytest

import av1gpu_amd as A
import refcheck
from test_gate_image import gate_planes


def post_params(lf=(0, 0, 0, 0), sharp=0, cdef=None):
    """frame parameters carrying loop filter levels / sharpness and, with cdef = (damping-3, y_pri, y_sec, uv_pri,
    uv_sec) or a list of such (cdef_bits = log2 of its length), CDEF"""
    fp = A._FrameParams()
    fp.version = 5
    for k in range(4):
        fp.loop_filter_level[k] = lf[k]
    fp.loop_filter_sharpness = sharp
    if cdef is not None:
        sets = cdef if isinstance(cdef, list) else [cdef]
        fp.enable_cdef = 1
        fp.cdef_damping_minus_3 = sets[0][0]
        fp.cdef_bits = {1: 0, 2: 1, 4: 2, 8: 3}[len(sets)]
        for i, (_, yp, ys, up, us) in enumerate(sets):
            fp.cdef_y_pri[i], fp.cdef_y_sec[i], fp.cdef_uv_pri[i], fp.cdef_uv_sec[i] = yp, ys, up, us
    return fp


CASES = [(200, 136, 3, 120, 8, 1, "img"), (211, 149, 0, 60, 8, 1, "gate"), (136, 72, 50, 160, 10, 0, "img")]
POSTS = [dict(lf=(24, 24, 24, 24)), dict(cdef=(0, 4, 2, 2, 1)), dict(lf=(40, 17, 9, 30), cdef=(3, 7, 4, 3, 0)),
         dict(lf=(63, 63, 63, 63), sharp=5, cdef=(2, 15, 1, 15, 4))]


@pytest.fixture(scope="module")
def records(oracle):
    out = {}
    for W, H, seed, q, bd, ssx, kind in CASES:
        planes = gate_planes() if kind == "gate" else A.synth_image(seed, W, H, ssx=ssx, bd=bd)
        out[(W, H)] = (planes, oracle.encode(planes, W, H, ssx=ssx, bd=bd, qindex=q, threads=4))
    return out


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d" % (c[0], c[1]))
@pytest.mark.parametrize("post", POSTS, ids=lambda p: "-".join(sorted(p)))
def test_reference_postfilter_is_dav1d_exact(ref, records, case, post):
    if not refcheck.dav1d_available():
        pytest.skip("dav1d (Pillow's libavif) is not available on this host")
    W, H, seed, q, bd, ssx, _ = case
    _, r = records[(W, H)]
    fp = post_params(**post)
    rec_pf, avif, _ = ref.encode(r, W, H, q, ssx=ssx, bd=bd, avif=True, post=fp)
    rec0, _, _ = ref.encode(r, W, H, q, ssx=ssx, bd=bd)
    d = refcheck.dav1d_decode_yuv(avif)
    for p in range(3):
        assert np.array_equal(d[p].astype(np.uint16), rec_pf[p].astype(np.uint16)), "reference writer, dav1d plane %d" % p
    assert any(not np.array_equal(rec_pf[p], rec0[p]) for p in range(3)), "the post-filters changed nothing"
    # this build's writer with the same parameters
    g = dict(width=W, height=H, base_q_idx=q, bit_depth=bd, blocks=r["blocks"], block_txb=r["block_txb"], txb=r["txb"],
             coeffs=r["coeffs"], tile_cols=1, tile_rows=1, chroma_format=0 if ssx else 1, fp=fp)
    d2 = refcheck.dav1d_decode_yuv(A.write_bitstream(g))
    for p in range(3):
        assert np.array_equal(d2[p].astype(np.uint16), rec_pf[p].astype(np.uint16)), "this writer, dav1d plane %d" % p
    parsed = ref.parse(A.write_bitstream(g, avif=False))["post"]
    for f in refcheck.Ref.PF_FIELDS:
        v = getattr(fp, f)
        want = list(v) if hasattr(v, "__len__") else v
        if f.startswith("cdef_") and f != "cdef_bits" and f != "cdef_damping_minus_3" and hasattr(v, "__len__"):
            n = 1 << fp.cdef_bits
            want, got = want[:n] if fp.enable_cdef else [], parsed[f][:n] if fp.enable_cdef else []
        else:
            got = parsed[f]
        assert got == want, f


def test_zero_levels_and_strengths_change_nothing(ref, records):
    W, H, seed, q, bd, ssx, _ = CASES[0]
    _, r = records[(W, H)]
    rec0, _, _ = ref.encode(r, W, H, q, ssx=ssx, bd=bd)
    rec1, _, _ = ref.encode(r, W, H, q, ssx=ssx, bd=bd, post=post_params(cdef=(0, 0, 0, 0, 0)))
    for p in range(3):
        assert np.array_equal(rec0[p], rec1[p])


def test_writer_codes_per_superblock_cdef_index(ref, records):
    """cdef_bits 2: every superblock's index coded with its first coded block; the reference parser reads the record back
    with the same indices and dav1d decodes it to the reference backend's reconstruction."""
    if not refcheck.dav1d_available():
        pytest.skip("dav1d (Pillow's libavif) is not available on this host")
    W, H, seed, q, bd, ssx, _ = CASES[0]
    _, r = records[(W, H)]
    blocks = r["blocks"].copy()
    sb = (blocks["mi_row"] >> 4) * ((W + 63) // 64) + (blocks["mi_col"] >> 4)
    blocks["cdef_idx"] = (sb * 3 + 1) % 4
    rr = dict(r, blocks=blocks)
    fp = post_params(lf=(20, 20, 20, 20), cdef=[(1, 2, 1, 1, 0), (1, 6, 2, 3, 1), (1, 11, 4, 0, 2), (1, 15, 0, 7, 4)])
    rec_pf, _, _ = ref.encode(rr, W, H, q, ssx=ssx, bd=bd, post=fp)
    g = dict(width=W, height=H, base_q_idx=q, bit_depth=bd, blocks=blocks, block_txb=r["block_txb"], txb=r["txb"],
             coeffs=r["coeffs"], tile_cols=1, tile_rows=1, chroma_format=0, fp=fp)
    parsed = ref.parse(A.write_bitstream(g, avif=False))
    coded = parsed["blocks"]["skip"] == 0
    assert np.array_equal(parsed["blocks"]["cdef_idx"][coded], blocks["cdef_idx"][coded])
    d = refcheck.dav1d_decode_yuv(A.write_bitstream(g))
    for p in range(3):
        assert np.array_equal(d[p], rec_pf[p]), p
