"""The oracle's open-loop analysis (screening -> top-K -> RD with tx-type search, k_analysis.cuh:496-689) against a
brute force built from pinned primitives: every candidate of every block of small frames is predicted with the
oracle's prediction and edge routines (both bit-exact with the reference's dsp.h), its SATD taken with a numpy 4x4
Hadamard, its mode bits priced from the spec's default CDFs, and the screening picks recomputed: the 13 modes, then
angle deltas +-1..3 around the two best directional modes, the K cheapest of all of them.  The oracle's candidate set
must be that set; its winner the cheapest of them under D + lambda (R + mode bits + overhead) with D and R from the
pinned tx-block evaluation, and its J and tx type the best over the winner's tx set.  Availability of the above-right
and below-left edges follows the decode order of the 4x4 units inside a superblock, computed here by walking the
quadtree (not by the oracle's bit-interleaving)."""
import ctypes as C

import numpy as np
import pytest

import av1gpu_amd as A
from test_oracle_pins import txg_table
from test_rate_kat import TXW, TXH, intra_set
from test_tables import _ours

T = _ours()
SHAPES = [s for s in range(22) if not 13 <= s <= 15]
BW = [4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 32, 64, 64, 64, 128, 128, 4, 16, 8, 32, 16, 64]
BH = [4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 64, 32, 64, 128, 64, 128, 16, 4, 32, 8, 64, 16]
TX_OF = {(TXW[t], TXH[t]): t for t in range(19)}
H4 = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]], np.int64)


def _decode_index():
    """Decode order of the 16 x 16 4x4 units of a superblock under recursive quadrant splitting."""
    idx = np.zeros((16, 16), np.int64)
    k = [0]

    def walk(x, y, n):
        if n == 1:
            idx[y, x] = k[0]
            k[0] += 1
            return
        h = n // 2
        for dy, dx in ((0, 0), (0, h), (h, 0), (h, h)):
            walk(x + dx, y + dy, h)
    walk(0, 0, 16)
    return idx


DEC = _decode_index()


def decoded_before(px, py, x, y, FW, FH):
    """Open-loop guess: the luma sample (px, py) is reconstructed before the block at (x, y) (one tile, frame
    FW x FH): rows of superblocks above, superblocks to the left in the same row, and inside the superblock the 4x4
    units earlier in decode order."""
    if px < 0 or py < 0 or px >= FW or py >= FH:
        return 0
    sbx, sby = x // 64 * 64, y // 64 * 64
    if py < sby:
        return 1
    if px < sbx:
        return int(py < sby + 64)
    if px >= sbx + 64 or py >= sby + 64:
        return 0
    return int(DEC[(py - sby) // 4, (px - sbx) // 4] < DEC[(y - sby) // 4, (x - sbx) // 4])


def mode_bits_tables():
    ym = np.array(T["Default_Intra_Frame_Y_Mode_Cdf"], np.float64).reshape(5, 5, 14)
    p = np.zeros(13)
    for m in range(13):
        p[m] = np.mean([(ym[a, l, m] - (ym[a, l, m - 1] if m else 0)) / 32768.0 for a in range(5) for l in range(5)])
    avg_y = np.float32(-np.log2(np.maximum(p, 1e-6)))
    ad = np.array(T["Default_Angle_Delta_Cdf"], np.float64).reshape(8, 8)
    ang = np.zeros((8, 7), np.float32)
    for m in range(8):
        for s in range(7):
            pr = (ad[m, s] - (ad[m, s - 1] if s else 0)) / 32768.0
            ang[m, s] = np.float32(-np.log2(max(pr, 1.0 / 32768.0)))
    return avg_y, ang


def fmaf(a, b, c):
    return np.float32(np.float64(np.float32(a)) * np.float64(np.float32(b)) + np.float64(np.float32(c)))


def has_chroma(bs, row4, col4):
    return not ((BW[bs] == 4 and col4 % 2 == 0) or (BH[bs] == 4 and row4 % 2 == 0))


def satd(src, pred):
    d = (src.astype(np.int64) - pred.astype(np.int64))
    h, w = d.shape
    t = 0
    for y in range(0, h, 4):
        for x in range(0, w, 4):
            t += int(np.abs(H4 @ d[y:y + 4, x:x + 4] @ H4.T).sum())
    return t


def brute_block(oracle, plane, FW, FH, bs, bx, by, q, bd, lam, lamS, avg_y, ang, avg_skip0, K=6):
    W, Hh = BW[bs], BH[bs]
    x, y = bx * W, by * Hh
    t = TX_OF[(W, Hh)]
    hl, ha = int(x > 0), int(y > 0)
    har, hbl = decoded_before(x + W, y - 1, x, y, FW, FH), decoded_before(x - 1, y + Hh, x, y, FW, FH)
    maxX, maxY = FW - 1, FH - 1
    a, l = np.zeros(160, np.uint16), np.zeros(160, np.uint16)
    p16 = np.ascontiguousarray(plane.astype(np.uint16))
    oracle.lib.orc_build_edges(p16.ctypes.data, p16.shape[1], x, y, t, hl, ha, har, hbl, maxX, maxY, bd, a.ctypes.data, l.ctypes.data)
    src = p16[y:y + Hh, x:x + W]

    def pred(m, d, ef):
        aa, ll = a.copy(), l.copy()
        out = np.zeros(4096, np.int32)
        oracle.lib.orc_predict(m, d, ef, 0, hl, ha, t, x, y, maxX, maxY, aa.ctypes.data, ll.ctypes.data, bd, out.ctypes.data)
        return out[:W * Hh].reshape(Hh, W)

    def bits(m, d):
        b = np.float32(avg_y[m])
        if bs >= 3 and 1 <= m <= 8:
            b = np.float32(b + ang[m - 1][d + 3])
        return b

    js = {}
    for m in range(13):
        js[(m, 0)] = fmaf(lamS, bits(m, 0), np.float32(satd(src, pred(m, 0, 0)) * 0.25))
    if bs >= 3:
        dirs = sorted(range(1, 9), key=lambda m: (js[(m, 0)], m))[:2]
        for m in dirs:
            for d in (-1, -2, -3, 1, 2, 3):
                js[(m, d)] = fmaf(lamS, bits(m, d), np.float32(satd(src, pred(m, d, 0)) * 0.25))
    order = sorted(js, key=lambda c: js[c])
    top = order[:K]
    tie_screen = len(order) > K and abs(float(js[order[K - 1]]) - float(js[order[K]])) <= 1e-5 * abs(float(js[order[K]]))
    overhead = np.float32((np.float32(1.5) if has_chroma(bs, y // 4, x // 4) else np.float32(0)) + np.float32(avg_skip0))
    res_cache = {}

    def rd(c, tt):
        if c not in res_cache:
            res_cache[c] = np.ascontiguousarray((src.astype(np.int32) - pred(c[0], c[1], 1)).astype(np.int16))
        res = res_cache[c]
        out = np.zeros(3, np.float32)
        qv, co = np.zeros(1024, np.int16), np.zeros(1024, np.float32)
        oracle.lib.orc_eval_txb(res.ctypes.data, t, tt, 0, q, bd, 0, 0, c[0], C.c_float(0.375), C.c_float(0.656),
                                out.ctypes.data, qv.ctypes.data, co.ctypes.data)
        return fmaf(lam, np.float32(np.float32(out[1] + bits(*c)) + overhead), out[0])

    jk = [rd(c, 0) for c in sorted(top, key=lambda c: order.index(c))]
    ranked = sorted(top, key=lambda c: order.index(c))
    win = ranked[int(np.argmin(jk))]
    srt = sorted(jk)
    tie_rd = len(srt) > 1 and abs(float(srt[0]) - float(srt[1])) <= 1e-5 * abs(float(srt[0]))
    best_j, best_tx = float(min(jk)), 0
    for tt in intra_set(t):
        if tt != 0:
            j = float(rd(win, tt))
            if j < best_j:
                best_j, best_tx = j, tt
    return dict(top=set(top), win=win, J=best_j, tx=best_tx, tie_screen=tie_screen, tie_rd=tie_rd)


@pytest.mark.parametrize("W,H,seed,q,bd", [(64, 64, 1, 120, 8), (40, 24, 2, 60, 8), (72, 64, 3, 180, 10)])
def test_analysis_equals_brute_force(oracle, W, H, seed, q, bd):
    planes = A.synth_image(2500 + seed, W, H, bd=bd)
    r = oracle.encode(planes, W, H, qindex=q, bd=bd, detail=True)
    padded, _, _ = A.pad_planes(planes, W, H)
    FW, FH = r["MiCols"] * 4, r["MiRows"] * 4
    txg, lamt, _ = txg_table(oracle)
    lam = np.float32(lamt[int(bd == 10)][q])
    lamS = np.float32(np.sqrt(np.float32(lam)))
    misc = [np.zeros(n, np.float32) for n in (304, 19, 13, 50, 1, 512, 512)]
    oracle.lib.orc_misc_tables(*(a.ctypes.data for a in misc))
    avg_y_oracle, avg_skip0 = misc[2], misc[4][0]
    avg_y, ang = mode_bits_tables()
    assert np.array_equal(avg_y, avg_y_oracle)
    checked = total = 0
    for bs in SHAPES:
        w, h = BW[bs], BH[bs]
        nbx, nby = (FW + w - 1) // w, (FH + h - 1) // h
        rec = r["blockres"][bs]
        for by in range(nby):
            for bx in range(nbx):
                if bx * w >= FW or by * h >= FH:
                    continue
                b = brute_block(oracle, padded[0], FW, FH, bs, bx, by, q, bd, lam, lamS, avg_y, ang, avg_skip0)
                total += 1
                got = rec[by * nbx + bx]
                cands = {(int(c) >> 3, (int(c) & 7) - 3) for c in got["cand"][:got["ncand"]]}
                if not b["tie_screen"]:
                    assert cands == b["top"], (bs, bx, by)
                    if not b["tie_rd"]:
                        c0 = int(got["cand"][0])
                        assert (c0 >> 3, (c0 & 7) - 3) == b["win"], (bs, bx, by)
                        assert abs(float(got["J"]) - b["J"]) <= 1e-5 * abs(b["J"]) + 1e-3, (bs, bx, by, got["J"], b["J"])
                        checked += 1
    assert checked >= 0.9 * total, (checked, total)   # near-ties (within 1e-5) are the only blocks left unchecked
