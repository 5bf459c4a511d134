"""GPU path (HIP kernels through the C ABI) against the oracle, bit for bit: every record array and the whole padded
reconstruction, over sizes (partial superblocks, 8x8, two-tile frames), qindex values, 8 and 10 bits, non-default
search knobs, and batches mixing sizes and qindex (each frame's result independent of the batch).  At full size the
GPU record is checked by the reference's CPU backend: its decoder-exact reconstruction equals the GPU's.  The search
internals -- every analysis record (J, candidates, tx type), every superblock's DP cost and its leaves -- equal the
oracle's bit for bit, so the decisions are made at the same RD costs."""
import numpy as np
import pytest

import av1gpu_amd as A
import refcheck

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("W,H,seed,q", [(64, 64, 1, 120), (128, 96, 2, 80), (200, 136, 3, 160), (8, 8, 4, 120),
                                        (72, 40, 5, 255), (264, 200, 6, 120), (512, 512, 7, 120), (130, 18, 8, 30)])
def test_gpu_equals_oracle(enc, oracle, W, H, seed, q):
    print("case", W, H, seed, q, flush=True)
    planes = A.synth_image(1000 + seed, W, H)
    g = enc.encode([A.DeviceFrame(planes, W, H, qindex=q)])[0][0]
    r = oracle.encode(planes, W, H, qindex=q, threads=8)
    _same(g, r)


def test_gpu_10bit_equals_oracle(enc, oracle):
    W, H = 136, 72
    planes = A.synth_image(1050, W, H, bd=10)
    g = enc.encode([A.DeviceFrame(planes, W, H, bd=10)])[0][0]
    r = oracle.encode(planes, W, H, bd=10, qindex=120, threads=8)
    _same(g, r)


def test_gpu_knobs_equal_oracle(oracle):
    """Non-default knobs: fewer candidates, the separate tx-type pass instead of txAllCL, uvK 2, no angle deltas."""
    P = A.GpParams(K=3, Krecheck=2, txAllCL=0, uvK=2, deltas=0, uvDeltas=0)
    e = A.GpEncoder().init(A.GpConfig(qindex=100, P=P))
    W, H = 96, 64
    planes = A.synth_image(1060, W, H)
    g = e.encode([A.DeviceFrame(planes, W, H)])[0][0]
    r = oracle.encode(planes, W, H, qindex=100, params=oracle.params(K=3, Krecheck=2, txAllCL=0, uvK=2, deltas=0, uvDeltas=0))
    _same(g, r)
    e.close()


def test_gpu_batch_is_independent_of_composition(enc):
    """Frames of different sizes and qindex in one call give the same records as alone (and in reversed order)."""
    specs = [(96, 64, 1070, 80), (200, 136, 1071, 120), (64, 130, 1072, 160)]
    frames = [A.DeviceFrame(A.synth_image(s, W, H), W, H, qindex=q) for W, H, s, q in specs]
    alone = [enc.encode([f])[0][0] for f in frames]
    together, t = enc.encode(frames)
    rev, _ = enc.encode(frames[::-1])
    for a, b, c in zip(alone, together, rev[::-1]):
        for x in (b, c):
            assert np.array_equal(a.blocks, x.blocks) and np.array_equal(a.txb, x.txb) and np.array_equal(a.coeffs, x.coeffs)
            for p in range(3):
                assert np.array_equal(a.recon[p], x.recon[p])
    assert t.blocks == sum(len(a.blocks) for a in alone)


def test_gpu_flat_frame_in_batch_equals_oracle(enc, oracle):
    """The host records' coefficients come down as nonzero masks and values: a flat frame at qindex 255 (no nonzero
    level at all) between textured frames whose coefficient counts end inside a mask span equals the oracle, each frame
    in the batch and alone."""
    W, H = 136, 72
    flat = [np.full((H, W), 128, np.uint8), np.full((H // 2, W // 2), 128, np.uint8), np.full((H // 2, W // 2), 128, np.uint8)]
    specs = [(A.synth_image(1090, 96, 40), 96, 40, 120), (flat, W, H, 255), (A.synth_image(1091, 72, 56), 72, 56, 60)]
    frames = [A.DeviceFrame(p, w, h, qindex=q) for p, w, h, q in specs]
    together, t = enc.encode(frames)
    assert int(np.count_nonzero(together[1].coeffs)) == 0 and len(together[1].coeffs) > 0
    assert any(len(g.coeffs) % 512 for g in together)
    for (p, w, h, q), f, g in zip(specs, frames, together):
        print("frame %dx%d q%d, %d coefficients" % (w, h, q, len(g.coeffs)), flush=True)
        _same(g, oracle.encode(p, w, h, qindex=q, threads=8))
        _same(enc.encode([f])[0][0], oracle.encode(p, w, h, qindex=q, threads=8))


def _decodes(ref, g, W, H, q):
    dec, avif, info = ref.encode(dict(blocks=g.blocks, block_txb=g.block_txb, txb=g.txb, coeffs=g.coeffs), W, H, q, avif=True)
    for p in range(3):
        assert np.array_equal(g.recon[p][:dec[p].shape[0], :dec[p].shape[1]], dec[p]), "plane %d" % p
    assert np.array_equal(info["eob"], g.txb["eob"])
    # this build's own writer: its stream parses back to the GPU record in the reference's parser
    parsed = ref.parse(A.write_bitstream(g, avif=False))
    for k in ("blocks", "block_txb", "txb", "coeffs"):
        assert np.array_equal(parsed[k], getattr(g, k)), k
    if refcheck.dav1d_available():
        for f in (avif, A.write_bitstream(g)):
            d = refcheck.dav1d_decode_yuv(f)
            for p in range(3):
                assert np.array_equal(d[p], dec[p]), "dav1d plane %d" % p


def test_gpu_record_decodes_to_gpu_reconstruction(enc, ref):
    W, H = 512, 512
    planes = A.synth_image(1000, W, H)
    g = enc.encode([A.DeviceFrame(planes, W, H)])[0][0]
    _decodes(ref, g, W, H, 120)


def test_gpu_full_size_record_decodes(enc, ref):
    """4000 x 3000 (two tiles, 2961 superblocks): the reference backend's reconstruction of the GPU record is the GPU's."""
    W, H = 4000, 3000
    planes = A.synth_image(1000, W, H)
    g, t = enc.encode([A.DeviceFrame(planes, W, H)])
    g = g[0]
    assert g.tile_rows == 2 and t.sbs == 63 * 47
    _decodes(ref, g, W, H, 120)


def test_gpu_schedules_agree(enc):
    """The persistent queue (every resident slot, and only 3 workgroups, which makes most superblocks wait on their
    neighbours; its one- and two-workgroups-per-CU instances) and one launch per wavefront step give identical records
    for a batch of mixed sizes and qindex."""
    specs = [(1000, 700, 1080, 120), (300, 200, 1081, 80), (64, 64, 1082, 160), (520, 130, 1083, 120)]
    frames = [A.DeviceFrame(A.synth_image(s, W, H), W, H, qindex=q) for W, H, s, q in specs]
    base, tq = enc.encode(frames)
    assert tq.schedule == 0 and tq.queueBlocks > 0
    for sched, qb, occ in ((1, 0, 0), (0, 3, 0), (0, 0, 1), (0, 0, 2), (0, 5, 2)):
        e = A.GpEncoder().init(A.GpConfig(qindex=120, schedule=sched, queueBlocks=qb, queueOcc=occ))
        print("schedule", sched, "queueBlocks", qb, "queueOcc", occ, flush=True)
        out, t = e.encode(frames)
        assert t.schedule == sched
        for a, b in zip(base, out):
            assert np.array_equal(a.blocks, b.blocks) and np.array_equal(a.txb, b.txb) and np.array_equal(a.coeffs, b.coeffs)
            for p in range(3):
                assert np.array_equal(a.recon[p], b.recon[p])
        e.close()


def same_search(g, r):
    """The GPU's search internals (GpEncoder.search_export) equal the oracle's (detail=True), bit for bit."""
    for s, a in g["blockres"].items():
        b = r["blockres"][s]
        assert len(a) == len(b), s
        assert np.array_equal(a["J"].view(np.uint32), b["J"].view(np.uint32)), ("J", s)
        assert np.array_equal(a["ncand"], b["ncand"]) and np.array_equal(a["tx"], b["tx"]), ("ncand/tx", s)
        idx = np.arange(6)[None, :] < a["ncand"][:, None]
        assert np.array_equal(np.where(idx, a["cand"], 0), np.where(idx, b["cand"], 0)), ("cand", s)
    assert np.array_equal(g["sbJ"].view(np.uint32), r["sbJ"].view(np.uint32)), "sbJ"
    assert np.array_equal(g["leaf_count"], r["leaf_count"]), "leaf counts"
    for k in range(len(g["leaf_count"])):
        n = int(g["leaf_count"][k])
        for f in ("bsize", "x4", "y4"):
            assert np.array_equal(g["leaves"][k][:n][f], r["leaves"][k][:n][f]), ("leaves", k)


@pytest.mark.parametrize("W,H,seed,q,bd", [(64, 64, 1, 120, 8), (200, 136, 3, 160, 8), (130, 18, 8, 30, 8), (136, 72, 50, 120, 10),
                                           (264, 200, 6, 80, 8)])
def test_gpu_search_costs_equal_oracle(oracle, W, H, seed, q, bd):
    print("search export", W, H, seed, q, bd, flush=True)
    e = A.GpEncoder().init(A.GpConfig(qindex=q))
    e.set_search_export(True)
    planes = A.synth_image(1000 + seed, W, H, bd=bd)
    g = e.encode([A.DeviceFrame(planes, W, H, bd=bd, qindex=q)])[0][0]
    ex = e.search_export(0)
    e.close()
    r = oracle.encode(planes, W, H, bd=bd, qindex=q, threads=8, detail=True)
    _same(g, r)
    same_search(ex, r)
