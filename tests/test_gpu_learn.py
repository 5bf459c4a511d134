"""The pruning network's kernels against the float64 NumPy restatement (tests/ln_ref.py), layer by layer: the MFMA
convolutions and the FC head (forward), the backward kernels (gradients), Adam training lowering the loss, and the
encoder integration (a network that expands every node and keeps every mode gives the exhaustive records bit for bit;
one that prunes gives a record the reference's backend decodes to the GPU reconstruction)."""
import numpy as np
import pytest

import av1gpu_amd as A
from av1gpu_amd import learn as Ln

import ln_ref as R

pytestmark = pytest.mark.gpu


def _data(n_img=2, seed=3000):
    tiles, qs = [], []
    for i in range(n_img):
        W, H = 256, 192
        y = A.synth_image(seed + i, W, H)[0]
        t = Ln.sb_tiles(y, W, H)
        tiles.append(t)
        qs.append(np.full(len(t), (80, 160)[i % 2], np.uint8))
    return np.concatenate(tiles), np.concatenate(qs)


def test_forward_matches_reference():
    p = Ln.init_params(11)
    tiles, q = _data()
    L = Ln.Learner(p)
    logits = L.forward(tiles, q)
    n = len(tiles)
    acts = [L.activations(k, n) for k in range(5)]
    x0 = R.input_tensor(tiles, q)
    assert np.array_equal(acts[0].reshape(x0.shape), x0)
    convs, w5, b5 = Ln.unpack(p)
    shapes = [(n, 64, 64, 2), (n, 32, 32, 32), (n, 16, 16, 32), (n, 8, 8, 32), (n, 4, 4, 32)]
    for l in range(4):
        inp = acts[l].reshape(shapes[l])
        ref = R.ref_conv(inp, convs[l][0], convs[l][1], l)
        got = acts[l + 1].reshape(shapes[l + 1])
        # FP32 accumulation in another order than float64: outputs agree to BF16 rounding (one step at most)
        assert np.allclose(got, ref, rtol=2 ** -7, atol=1e-4), "layer %d" % (l + 1)
        assert np.mean(got != ref) < 0.02
    zr = R.ref_fc(acts[4].reshape(n, -1), w5, b5)
    assert np.allclose(logits, zr, rtol=1e-4, atol=1e-4)
    L.close()


def test_gradient_matches_reference():
    p = Ln.init_params(12)
    tiles, q = _data()
    rng = np.random.default_rng(5)
    lab = (rng.uniform(size=(len(tiles), Ln.N_OUT)) < 0.4).astype(np.float32)
    msk = (rng.uniform(size=(len(tiles), Ln.N_OUT)) < 0.9).astype(np.float32)
    L = Ln.Learner(p)
    L.set_data(tiles, q, lab, msk)
    loss, g = L.gradient(0, len(tiles))
    acts = [L.activations(k, len(tiles)) for k in range(5)]
    shapes = [(64, 64, 2), (32, 32, 32), (16, 16, 32), (8, 8, 32), (4, 4, 32)]
    acts = [a.reshape((len(tiles),) + s) for a, s in zip(acts, shapes)]
    lr, gr = R.ref_grads(p, acts, lab, msk)
    assert abs(loss - lr) < 1e-4 * max(1.0, abs(lr))
    gc, gw5, gb5 = Ln.unpack(g.astype(np.float64))
    rc, rw5, rb5 = Ln.unpack(gr)
    for a, b, name in [(gw5, rw5, "fc W"), (gb5, rb5, "fc b")] + [(gc[l][0], rc[l][0], "conv%d W" % l) for l in range(4)] + \
            [(gc[l][1], rc[l][1], "conv%d b" % l) for l in range(4)]:
        assert np.max(np.abs(a - b)) <= 1e-3 * np.max(np.abs(b)) + 1e-7, name
    L.close()


def test_training_lowers_the_loss():
    tiles, q = _data(4)
    rng = np.random.default_rng(6)
    # a learnable target: the node outputs follow the tile's variance, the modes a fixed pattern
    var = tiles.reshape(len(tiles), -1).astype(np.float64).std(1)
    lab = np.zeros((len(tiles), Ln.N_OUT), np.float32)
    lab[:, :Ln.MODE0] = (var > np.median(var))[:, None]
    lab[:, Ln.MODE0:] = rng.dirichlet(np.ones(13))
    L = Ln.Learner(Ln.init_params(13))
    L.set_data(tiles, q, lab, np.ones_like(lab))
    first = L.step(0, len(tiles), 3e-3)
    for _ in range(60):
        last = L.step(0, len(tiles), 3e-3)
    assert last < 0.8 * first
    L.close()


def _params_bias(split_bias, mode_bias, own_bias=20.0, four_bias=20.0):
    p = Ln.init_params(1)
    convs, w5, b5 = Ln.unpack(p)
    w5[:] = 0.0
    b5[:Ln.N_SPLIT] = split_bias
    b5[Ln.OWN0:Ln.FOUR0] = own_bias
    b5[Ln.FOUR0:Ln.MODE0] = four_bias
    b5[Ln.MODE0:] = mode_bias
    return p


def test_shortlist_that_keeps_everything_is_exhaustive():
    W, H = 264, 200
    f = A.DeviceFrame(A.synth_image(1006, W, H), W, H, qindex=120)
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
    base = enc.encode([f])[0][0]
    Ln.set_shortlist(enc, _params_bias(20.0, np.zeros(13)), 0.8)   # every node expanded, uniform modes: no shortlist
    got = enc.encode([f])[0][0]
    for k in ("blocks", "txb", "coeffs"):
        assert np.array_equal(getattr(base, k), getattr(got, k)), k
    for p in range(3):
        assert np.array_equal(base.recon[p], got.recon[p])
    enc.close()
    f.close()


def test_pruned_record_decodes(ref):
    W, H = 264, 200
    f = A.DeviceFrame(A.synth_image(1007, W, H), W, H, qindex=120)
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
    base = enc.encode([f])[0][0]
    mb = np.full(13, -10.0)
    mb[[0, 1, 2, 9, 10, 11, 12, 3]] = 10.0                          # 8 modes kept
    Ln.set_shortlist(enc, _params_bias(-20.0, mb), 0.8)             # nothing expanded below 64x64 where not forced
    g = enc.encode([f])[0][0]
    assert len(g.blocks) < len(base.blocks) and set(np.unique(g.blocks["y_mode"])) <= {0, 1, 2, 3, 9, 10, 11, 12}
    dec, _, _ = ref.encode(dict(blocks=g.blocks, block_txb=g.block_txb, txb=g.txb, coeffs=g.coeffs), W, H, 120)
    for p in range(3):
        assert np.array_equal(g.recon[p][:dec[p].shape[0], :dec[p].shape[1]], dec[p])
    enc.close()
    f.close()


def test_network_chunks_and_batching_do_not_change_records(monkeypatch):
    """The masks of a superblock depend on its own pixels only: a batch of mixed-size images run through the network 7
    superblocks at a time (chunks straddle images) gives each image the record it gets alone in one chunk."""
    p = Ln.init_params(21)
    convs, w5, b5 = Ln.unpack(p)
    w5 *= 4.0                                                       # spread the logits so the masks vary per superblock
    sizes = [(264, 200), (136, 72), (328, 264), (64, 64)]
    frames = [A.DeviceFrame(A.synth_image(1100 + i, W, H), W, H, qindex=(80, 120, 160, 120)[i]) for i, (W, H) in enumerate(sizes)]
    solo = A.GpEncoder().init(A.GpConfig(qindex=120))
    Ln.set_shortlist(solo, p, 0.8)
    alone = [solo.encode([f])[0][0] for f in frames]
    monkeypatch.setenv("AV1MI_LN_CHUNK", "7")
    batch = A.GpEncoder().init(A.GpConfig(qindex=120))
    Ln.set_shortlist(batch, p, 0.8)
    out = batch.encode(frames)[0]
    full = A.GpEncoder().init(A.GpConfig(qindex=120))
    exh = full.encode(frames)[0]
    pruned_any = False
    for a, b, e in zip(alone, out, exh):
        for k in ("blocks", "txb", "coeffs"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
        pruned_any |= len(a.blocks) != len(e.blocks) or not np.array_equal(a.blocks, e.blocks)
    assert pruned_any
    for e in (solo, batch, full):
        e.close()
    for f in frames:
        f.close()


def test_records_obey_the_restated_decisions():
    """Every block of a pruned record lies in a node the decisions (ln_ref.decide over the network's logits) expand, is
    one of the own (or 4-way) shapes of a node that keeps them, and uses a luma mode of its superblock's mask; a
    confidence where the masks are mixed is used (some nodes pruned, some superblocks shortlisted)."""
    p = Ln.init_params(22)
    convs, w5, b5 = Ln.unpack(p)
    w5 *= 4.0
    b5[Ln.MODE0:] = np.linspace(6.0, -6.0, 13)              # a few modes likely, the rest not: shortlists appear
    W, H = 328, 264
    planes = A.synth_image(1200, W, H)
    L = Ln.Learner(p)
    tiles = Ln.sb_tiles(planes[0], W, H)
    logits = L.forward(tiles, np.full(len(tiles), 120, np.uint8))
    L.close()
    f = A.DeviceFrame(planes, W, H, qindex=120)
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
    conf = 0.8
    Ln.set_shortlist(enc, p, conf)
    g = enc.encode([f])[0][0]
    expand, own, four, mmask = R.decide(logits, W, H, conf)
    assert 0 < expand[:, 1:].mean() < 1 and 0 < own.mean() < 1 and 0 < four.mean() < 1 and (mmask != 0x1FFF).any()
    sbc = (2 * ((W + 7) >> 3) + 15) >> 4
    w4 = np.array([1 << v for v in Ln._MI_W_LOG2])
    h4 = np.array([1 << v for v in Ln._MI_H_LOG2])
    for b in g.blocks:
        x, y, s = 4 * int(b["mi_col"]), 4 * int(b["mi_row"]), int(b["bsize"])
        sb = (y // 64) * sbc + x // 64
        assert (mmask[sb] >> int(b["y_mode"])) & 1
        M = 4 * max(w4[s], h4[s])
        lvl = {64: 0, 32: 1, 16: 2, 8: 3}
        if M < 64:
            n = 2 * M
            node = Ln.LEVEL_BASE[lvl[n]] + ((y % 64) // n) * (64 // n) + (x % 64) // n
            assert expand[sb, node], (x, y, s)
        if M > 4:
            node = Ln.LEVEL_BASE[lvl[M]] + ((y % 64) // M) * (64 // M) + (x % 64) // M
            assert (four if max(w4[s], h4[s]) == 4 * min(w4[s], h4[s]) else own)[sb, node], (x, y, s)
    enc.close()
    f.close()


def test_dp_export_is_consistent_with_the_records():
    """The partition DP's exported relative costs (training labels) change no record, are never negative, and agree
    with the choices at each superblock's root: one coded as a single 64x64 leaf chose NONE, so dropping the root's
    children or its 4-way shapes costs 0; one without any leaf of side 64 chose SPLIT, so dropping the root's own or
    4-way shapes costs 0; one with a 64x16 or 16x64 leaf chose a 4-way partition, so dropping the root's children or
    own shapes costs 0."""
    W, H = 328, 264
    f = A.DeviceFrame(A.synth_image(1300, W, H), W, H, qindex=255)
    e = A.GpEncoder().init(A.GpConfig(qindex=120))
    base = e.encode([f])[0][0]
    rel = Ln.dp_export(e, True)
    assert rel.size == 0
    g = e.encode([f])[0][0]
    rel = Ln.dp_export(e, None)
    for k in ("blocks", "txb", "coeffs"):
        assert np.array_equal(getattr(base, k), getattr(g, k)), k
    sbc = (2 * ((W + 7) >> 3) + 15) >> 4
    assert rel.shape == ((2 * ((H + 7) >> 3) + 15) // 16 * sbc, 255) and (rel > -1e-6).all()
    w4 = np.array([1 << v for v in Ln._MI_W_LOG2])
    h4 = np.array([1 << v for v in Ln._MI_H_LOG2])
    b = g.blocks
    sb = (b["mi_row"].astype(int) // 16) * sbc + b["mi_col"].astype(int) // 16
    side = 4 * np.maximum(w4[b["bsize"]], h4[b["bsize"]])
    none64 = checked = 0
    for s in range(len(rel)):
        r0, c0 = (s // sbc) * 64, (s % sbc) * 64
        if r0 + 64 > 4 * 2 * ((H + 7) >> 3) or c0 + 64 > 4 * 2 * ((W + 7) >> 3):
            continue                                     # edge superblocks: forced partitions
        bs = b["bsize"][sb == s]
        if len(bs) == 1 and bs[0] == 12:                 # PARTITION_NONE at the root
            assert rel[s, 0] == 0.0 and rel[s, 170] == 0.0
            none64 += 1
        if not (side[sb == s] == 64).any():              # SPLIT at the root
            assert rel[s, 85] == 0.0 and rel[s, 170] == 0.0
        if np.isin(bs, (20, 21)).any():                  # HORZ_4 / VERT_4 at the root
            assert rel[s, 0] == 0.0 and rel[s, 85] == 0.0
        checked += 1
    assert checked > 0
    e.close()
    f.close()
