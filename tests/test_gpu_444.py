"""4:4:4 input (chroma at full resolution, 8 and 10 bits) on the GPU path against the oracle,
bit for bit, including blocks with a 64-sample side whose chroma holds 2 or 4 tx blocks (the high-qindex cases), frame
edges, a batch of mixed sizes, the learned pruning, and the reference's backend / dav1d decoding the records."""
import numpy as np
import pytest

import av1gpu_amd as A
import refcheck
from av1gpu_amd import learn

pytestmark = pytest.mark.gpu

BIG = [10, 11, 12, 20, 21]   # 32x64, 64x32, 64x64, 64x16, 16x64


@pytest.fixture(scope="module")
def enc():
    e = A.GpEncoder().init(A.GpConfig(qindex=120))
    yield e
    e.close()


def _same(g, r):
    for k, a in (("blocks", g.blocks), ("block_txb", g.block_txb), ("txb", g.txb), ("coeffs", g.coeffs)):
        assert len(a) == len(r[k]) and np.array_equal(a, r[k]), k
    for p in range(3):
        assert np.array_equal(g.recon[p].astype(np.uint16), r["recon"][p]), "recon plane %d" % p


@pytest.mark.parametrize("W,H,seed,q,bd", [(200, 136, 1, 120, 8), (256, 256, 11, 255, 8), (200, 264, 13, 240, 10), (130, 18, 4, 200, 8),
                                           (136, 72, 3, 160, 10), (328, 200, 12, 255, 8)])
def test_gpu_444_equals_oracle(enc, oracle, W, H, seed, q, bd):
    print("case 444", W, H, seed, q, bd, flush=True)
    planes = A.synth_image(seed, W, H, ssx=0, bd=bd)
    g = enc.encode([A.DeviceFrame(planes, W, H, bd=bd, qindex=q, yuv444=True)])[0][0]
    assert g.chroma_format == 1 and g.recon[1].shape == g.recon[0].shape
    r = oracle.encode(planes, W, H, ssx=0, bd=bd, qindex=q, threads=8)
    if q >= 240:
        assert np.isin(r["blocks"]["bsize"], BIG).any()
    _same(g, r)


def test_gpu_444_batch_and_decodes(enc, ref):
    """A 4:4:4 batch of mixed sizes and qindex: each record equals its image alone and the two-workgroups-per-CU closed
    loop's; the reference's backend and dav1d decode each to the GPU reconstruction."""
    specs = [(264, 200, 21, 120), (96, 64, 22, 255), (200, 136, 23, 80)]
    frames = [A.DeviceFrame(A.synth_image(s, W, H, ssx=0), W, H, qindex=q, yuv444=True) for W, H, s, q in specs]
    alone = [enc.encode([f])[0][0] for f in frames]
    together = enc.encode(frames)[0]
    e2 = A.GpEncoder().init(A.GpConfig(qindex=120, queueOcc=2))   # the two-workgroups-per-CU closed loop
    other = e2.encode(frames)[0]
    e2.close()
    for a, b, c, (W, H, s, q) in zip(alone, together, other, specs):
        for k in ("blocks", "txb", "coeffs"):
            assert np.array_equal(getattr(a, k), getattr(b, k)) and np.array_equal(getattr(a, k), getattr(c, k)), k
        for p in range(3):
            assert np.array_equal(b.recon[p], c.recon[p])
        dec, _, _ = ref.encode(dict(blocks=b.blocks, block_txb=b.block_txb, txb=b.txb, coeffs=b.coeffs), W, H, q, ssx=0)
        for p in range(3):
            assert np.array_equal(b.recon[p][:dec[p].shape[0], :dec[p].shape[1]], dec[p])
        if refcheck.dav1d_available():
            d = refcheck.dav1d_decode_yuv(A.write_bitstream(b))
            for p in range(3):
                assert np.array_equal(d[p].astype(np.uint16), b.recon[p][:d[p].shape[0], :d[p].shape[1]].astype(np.uint16))
    for f in frames:
        f.close()


def test_gpu_mixed_chroma_formats_are_refused(enc):
    a = A.DeviceFrame(A.synth_image(31, 64, 64), 64, 64)
    b = A.DeviceFrame(A.synth_image(32, 64, 64, ssx=0), 64, 64, yuv444=True)
    with pytest.raises(RuntimeError, match="chroma formats"):
        enc.encode([a, b])
    a.close()
    b.close()


def test_gpu_444_10bit_learned_record_decodes(ref):
    """The batched+learned path on 10-bit 4:4:4: the pruned record decodes to the GPU reconstruction."""
    W, H = 328, 264
    planes = A.synth_image(41, W, H, ssx=0, bd=10)
    e = A.GpEncoder().init(A.GpConfig(qindex=120))
    learn.set_shortlist(e, learn.load_params(), 0.8)
    g = e.encode([A.DeviceFrame(planes, W, H, bd=10, qindex=120, yuv444=True)])[0][0]
    dec, _, _ = ref.encode(dict(blocks=g.blocks, block_txb=g.block_txb, txb=g.txb, coeffs=g.coeffs), W, H, 120, ssx=0, bd=10)
    for p in range(3):
        assert np.array_equal(g.recon[p][:dec[p].shape[0], :dec[p].shape[1]], dec[p])
    e.close()
