"""The oracle's integer restatement (oracle/orc_dsp.h) against the reference's own header-only DSP
(av1core/include/av1g/dsp.h, compiled from the reference sources into oracle/_ref/): dequantization + inverse
transforms of every size and type, intra prediction of every mode / delta / size with and without the edge filter,
edge building, tile layout.  Seeded random inputs, bit-exact."""
import ctypes as C

import numpy as np

TX_W = [4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64]
TX_H = [4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16]
# intra tx types allowed per tx size (sets DCT-only / INTRA_1 / INTRA_2)
SET1 = [9, 0, 10, 11, 3, 1, 2]
SET2 = [9, 0, 3, 1, 2]


def types_of(t):
    sq_up = max(TX_W[t], TX_H[t])
    if sq_up >= 32:
        return [0]
    if min(TX_W[t], TX_H[t]) == 16:
        return SET2
    return SET1


def test_dequant_inverse_matches_reference(oracle, ref):
    rng = np.random.default_rng(0x123456789ABCDEF)
    for bd in (8, 10):
        for t in range(19):
            tw, th = min(TX_W[t], 32), min(TX_H[t], 32)
            for ty in range(16) if max(TX_W[t], TX_H[t]) <= 16 else [0]:
                for trial in range(6):
                    q = np.zeros(tw * th, np.int16)
                    k = rng.integers(1, max(2, tw * th // (1 + trial)))
                    idx = rng.choice(tw * th, size=k, replace=False)
                    q[idx] = rng.integers(-60 * (trial + 1), 60 * (trial + 1), size=k)
                    dcq, acq = (int(v) for v in rng.integers(4, 400 if bd == 8 else 1600, size=2))
                    r1, r2 = np.zeros(4096, np.int32), np.zeros(4096, np.int32)
                    e1 = oracle.lib.orc_dequant_inverse(q.ctypes.data, t, ty, dcq, acq, bd, r1.ctypes.data)
                    e2 = ref.lib.ref_dequant_inverse(q.ctypes.data, t, ty, dcq, acq, bd, r2.ctypes.data)
                    n = TX_W[t] * TX_H[t]
                    assert e1 == e2, (bd, t, ty, trial)
                    assert np.array_equal(r1[:n], r2[:n]), (bd, t, ty, trial)


def test_prediction_matches_reference(oracle, ref):
    rng = np.random.default_rng(0x9E3779B97F4A7C15)
    EDGE_LEN = 160
    for bd in (8, 10):
        for t in range(19):
            for mode in range(13):
                for delta in ([-3, -1, 0, 2, 3] if 1 <= mode <= 8 else [0]):
                    for ef, ft in ((0, 0), (1, 0), (1, 1)):
                        hl, ha = (int(v) for v in rng.integers(0, 2, size=2))
                        a = rng.integers(0, 1 << bd, size=EDGE_LEN).astype(np.uint16)
                        l = rng.integers(0, 1 << bd, size=EDGE_LEN).astype(np.uint16)
                        x, y = int(rng.integers(0, 64)) * 4, int(rng.integers(0, 64)) * 4
                        maxX, maxY = x + int(rng.integers(1, 130)), y + int(rng.integers(1, 130))
                        p1, p2 = np.zeros(4096, np.int32), np.zeros(4096, np.int32)
                        a2, l2 = a.copy(), l.copy()
                        oracle.lib.orc_predict(mode, delta, ef, ft, hl, ha, t, x, y, maxX, maxY, a.ctypes.data, l.ctypes.data, bd,
                                               p1.ctypes.data)
                        ref.lib.ref_predict(mode, delta, ef, ft, hl, ha, t, x, y, maxX, maxY, a2.ctypes.data, l2.ctypes.data, bd,
                                            p2.ctypes.data)
                        n = TX_W[t] * TX_H[t]
                        assert np.array_equal(p1[:n], p2[:n]), (bd, t, mode, delta, ef, ft)


def test_build_edges_matches_reference(oracle, ref):
    rng = np.random.default_rng(7)
    plane = rng.integers(0, 1024, size=(256, 256)).astype(np.uint16)
    for trial in range(400):
        t = int(rng.integers(0, 19))
        x, y = int(rng.integers(1, 40)) * 4, int(rng.integers(1, 40)) * 4
        hl, ha, har, hbl = (int(v) for v in rng.integers(0, 2, size=4))
        maxX, maxY = min(255, x + int(rng.integers(0, 140))), min(255, y + int(rng.integers(0, 140)))
        a1, l1, a2, l2 = (np.zeros(160, np.uint16) for _ in range(4))
        oracle.lib.orc_build_edges(plane.ctypes.data, 256, x, y, t, hl, ha, har, hbl, maxX, maxY, 10, a1.ctypes.data, l1.ctypes.data)
        ref.lib.ref_build_edges(plane.ctypes.data, 256, x, y, t, hl, ha, har, hbl, maxX, maxY, 10, a2.ctypes.data, l2.ctypes.data)
        n = TX_W[t] + TX_H[t]
        assert np.array_equal(a1[15:16 + n], a2[15:16 + n]) and np.array_equal(l1[15:16 + n], l2[15:16 + n]), trial


def test_tile_layout_matches_reference(oracle, ref):
    for W, H in ((512, 512), (4000, 3000), (4096, 2304), (4100, 2304), (8192, 1024), (200, 136), (16384, 8), (7000, 5000)):
        cols, rows = C.c_int(), C.c_int()
        cs, rs = (C.c_int * 65)(), (C.c_int * 65)()
        assert oracle.lib.orc_tile_layout(W, H, C.byref(cols), C.byref(rows), cs, rs) == 0
        rc, rr = ref.tile_layout(W, H)
        assert list(cs[:cols.value + 1]) == rc and list(rs[:rows.value + 1]) == rr, (W, H)
