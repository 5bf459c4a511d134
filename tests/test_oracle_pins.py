"""Checks of the oracle's float search against what the mathematics fixes, independent of its own code:

* the dead-zone quantiser and distortion of every evaluated tx block, recomputed from its forward-transform outputs in
  numpy (float32 quantiser arithmetic as defined, float64 distortion): levels bit for bit, D to float rounding; the
  same recomputation with a shifted rounding offset disagrees (the check has teeth);
* the forward transform as the inverse of the exact inverse: the decoder's inverse of one dequantised coefficient
  (the reference-pinned integer path), placed into the block with the spec's flips, comes back from the forward
  transform as that coefficient and nearly nothing else, for every tx size and 2-D tx type (flips, the row scale
  and the 64-point truncation included);
* the partition DP against brute force: every partition tree of 8x8 and 16x16 nodes enumerated from the AV1 partition
  semantics, larger nodes by optimal substructure over those, in float64 from the analysis records; the oracle's
  superblock costs and leaves equal the minimum tree's (frames with partial superblocks, and restricted shape /
  partition sets); plus a full enumeration of every tree of small frames.
"""
import numpy as np
import pytest

from test_rate_kat import TXW, TXH, intra_set
from test_tables import _ours

T = _ours()
ALL_TYPES = list(range(16))


def q_steps(qindex, bd):
    k = (bd - 8) >> 1
    return T["Dc_Qlookup"][k * 256 + qindex], T["Ac_Qlookup"][k * 256 + qindex]


def dq_denom(t):
    area = TXW[t] * TXH[t]
    return 4 if area > 1024 and max(TXW[t], TXH[t]) == 64 and min(TXW[t], TXH[t]) >= 32 else \
        2 if area >= 512 and area > 256 and not (TXW[t] * TXH[t] == 512 and max(TXW[t], TXH[t]) == 64 and False) else 1


def spec_dq_denom(t):
    # spec get_dq_denom: 2 for 32x32, 16x32, 32x16, 16x64, 64x16; 4 for 64x64, 32x64, 64x32; else 1
    if t in (4, 11, 12):
        return 4
    if t in (3, 9, 10, 17, 18):
        return 2
    return 1


def dequant(levels, dc, ac, den, bd):
    """spec 7.12.3: magnitude * step, & 0xFFFFFF, >> log2(den), sign, clamp to the coefficient range."""
    step = np.full(levels.shape, ac, np.int64)
    step[0] = dc
    m = (np.abs(levels.astype(np.int64)) * step) & 0xFFFFFF
    m >>= {1: 0, 2: 1, 4: 2}[den]
    v = np.where(levels < 0, -m, m)
    lim = 1 << (7 + bd)
    return np.clip(v, -lim, lim - 1)


def eval_txb(oracle, res, t, tx_type, plane, qindex, bd, rnd=0.375, zbin=0.656):
    import ctypes as C
    W, H = TXW[t], TXH[t]
    tw, th = min(W, 32), min(H, 32)
    res = np.ascontiguousarray(res, dtype=np.int16)
    out = np.zeros(3, np.float32)
    q = np.zeros(tw * th, np.int16)
    coef = np.zeros(tw * th, np.float32)
    oracle.lib.orc_eval_txb(res.ctypes.data, t, tx_type, plane, qindex, bd, 0, 0, 0, C.c_float(rnd), C.c_float(zbin),
                            out.ctypes.data, q.ctypes.data, coef.ctypes.data)
    return out, q, coef


def txg_table(oracle):
    import ctypes as C
    txg, sc, ay, ap, a0, lam, las = (np.zeros(19 * 16, np.float32), np.zeros(19, np.float32), np.zeros(13, np.float32),
                                     np.zeros(50, np.float32), np.zeros(1, np.float32), np.zeros(512, np.float32),
                                     np.zeros(512, np.float32))
    oracle.lib.orc_misc_tables(*(a.ctypes.data for a in (txg, sc, ay, ap, a0, lam, las)))
    return txg.reshape(19, 16), lam.reshape(2, 256), ap.reshape(5, 10)


@pytest.mark.parametrize("t", range(19))
def test_quantizer_and_distortion_closed_forms(oracle, t):
    rng = np.random.default_rng(77 + t)
    txg, _, _ = txg_table(oracle)
    W, H = TXW[t], TXH[t]
    den = spec_dq_denom(t)
    for qindex, bd in ((20, 8), (120, 8), (255, 8), (90, 10)):
        dc, ac = q_steps(qindex, bd)
        for tx_type in intra_set(t):
            amp = (255 if bd == 8 else 1023) * rng.choice([0.05, 0.3, 1.0])
            res = np.clip(np.round(rng.normal(0, amp / 3, (H, W))), -(1 << bd) + 1, (1 << bd) - 1).astype(np.int16)
            out, q, coef = eval_txb(oracle, res, t, tx_type, 0, qindex, bd)
            v = coef.astype(np.float32)
            step = np.full(v.shape, ac, np.float32)
            step[0] = dc
            a = (np.abs(v) * np.float32(den)) / step                      # float32 IEEE ops
            lvl = np.where(a < np.float32(0.656), 0, np.floor(a + np.float32(0.375)).astype(np.int64))
            lvl = np.minimum(lvl, 16383)
            exp = np.where(v < 0, -lvl, lvl)
            assert np.array_equal(exp, q.astype(np.int64))
            assert np.abs(q.astype(np.int64)).max(initial=0) <= 16383
            eob = int(out[2])
            assert (eob == 0) == (not q.any()) and np.count_nonzero(q) <= eob <= len(q)
            dq = dequant(q, dc, ac, den, bd).astype(np.float64)
            dist = float(txg[t, tx_type]) * np.sum((v.astype(np.float64) - dq) ** 2)
            if W == 64 or H == 64:   # the dropped coefficients' energy: the residual energy less the kept energy, >= 0
                kept = float(txg[t, tx_type]) * np.sum(v.astype(np.float64) ** 2)
                dist += max(0.0, float(np.sum(res.astype(np.float64) ** 2)) - kept)
            assert abs(out[0] - dist) <= 1e-4 * dist + 1e-3, (qindex, tx_type, out[0], dist)
            # a different rounding offset in the recomputation changes some level of a busy block (teeth)
            if np.count_nonzero(q) > 8:
                alt = np.where(a < np.float32(0.656), 0, np.floor(a + np.float32(0.5)).astype(np.int64))
                assert not np.array_equal(np.where(v < 0, -np.minimum(alt, 16383), np.minimum(alt, 16383)), exp) or \
                    not np.any((a - np.floor(a) >= 0.5) & (a - np.floor(a) < 0.625) & (a >= 0.656))


def _types_2d(t):
    W, H = TXW[t], TXH[t]
    if max(W, H) == 64:
        return [0]
    if max(W, H) == 32:
        return [0, 9]
    return ALL_TYPES


def _flip(tx_type):
    # spec 7.13.3: FLIPADST in the column (vertical) kind flips up-down, in the row kind left-right
    ud = tx_type in (4, 6, 8, 14)
    lr = tx_type in (5, 6, 7, 15)
    return ud, lr


@pytest.mark.parametrize("t", range(19))
def test_forward_inverts_the_exact_inverse(oracle, t):
    rng = np.random.default_rng(500 + t)
    W, H = TXW[t], TXH[t]
    tw, th = min(W, 32), min(H, 32)
    qindex, bd = 40, 10
    dc, ac = q_steps(qindex, bd)
    den = spec_dq_denom(t)
    for tx_type in _types_2d(t):
        ud, lr = _flip(tx_type)
        for pos in [0, tw * th - 1] + list(rng.integers(0, tw * th, 3)):
            lv = np.zeros(tw * th, np.int16)
            lv[pos] = int(rng.choice([-1, 1]) * rng.integers(20, 60))
            res = np.zeros(W * H, np.int32)
            err = oracle.lib.orc_dequant_inverse(lv.ctypes.data, t, tx_type, dc, ac, bd, res.ctypes.data)
            assert err == 0
            r = res.reshape(H, W)
            r = r[::-1] if ud else r
            r = r[:, ::-1] if lr else r
            _, q, coef = eval_txb(oracle, np.ascontiguousarray(r), t, tx_type, 0, qindex, bd)
            dqv = float(dequant(lv, dc, ac, den, bd)[pos])
            tol = 0.03 * abs(dqv) + 6
            assert abs(coef[pos] - dqv) <= tol, (tx_type, pos, coef[pos], dqv)
            others = np.delete(coef, pos)
            assert np.abs(others).max(initial=0) <= tol, (tx_type, pos, np.abs(others).max(), dqv)


# ---------------------------------------------------------------------------------------------------------------
# partition DP vs brute force
NONE, HORZ, VERT, SPLIT, HORZ_A, HORZ_B, VERT_A, VERT_B, HORZ_4, VERT_4 = range(10)
BS = {(4, 4): 0, (4, 8): 1, (8, 4): 2, (8, 8): 3, (8, 16): 4, (16, 8): 5, (16, 16): 6, (16, 32): 7, (32, 16): 8,
      (32, 32): 9, (32, 64): 10, (64, 32): 11, (64, 64): 12, (4, 16): 16, (16, 4): 17, (8, 32): 18, (32, 8): 19,
      (16, 64): 20, (64, 16): 21}


def partition_blocks(p, n, x, y):
    """AV1 partition semantics (spec 5.11.4 decode_partition): the blocks (w, h, x, y) or sub-nodes ('node', n/2, x, y)
    a partition of an n x n node at (x, y) produces, in decode order."""
    h, q = n // 2, n // 4
    if p == NONE:
        return [(n, n, x, y)]
    if p == HORZ:
        return [(n, h, x, y), (n, h, x, y + h)]
    if p == VERT:
        return [(h, n, x, y), (h, n, x + h, y)]
    if p == SPLIT:
        if n == 8:
            return [(4, 4, x, y), (4, 4, x + 4, y), (4, 4, x, y + 4), (4, 4, x + 4, y + 4)]
        return [("node", h, x, y), ("node", h, x + h, y), ("node", h, x, y + h), ("node", h, x + h, y + h)]
    if p == HORZ_A:
        return [(h, h, x, y), (h, h, x + h, y), (n, h, x, y + h)]
    if p == HORZ_B:
        return [(n, h, x, y), (h, h, x, y + h), (h, h, x + h, y + h)]
    if p == VERT_A:
        return [(h, h, x, y), (h, h, x, y + h), (h, n, x + h, y)]
    if p == VERT_B:
        return [(h, n, x, y), (h, h, x + h, y), (h, h, x + h, y + h)]
    if p == HORZ_4:
        return [(n, q, x, y + i * q) for i in range(4)]
    return [(q, n, x + i * q, y) for i in range(4)]


class Brute:
    """Minimum-cost partition trees from the analysis records, in float64.  Cost of a tree = sum over its nodes of
    lambda * (partition bits) + the J of its blocks (NONE's J times splitBias); blocks outside the coded frame cost 0
    and are not coded; a node past the frame's bottom / right edge may only split or take the one partition the
    frame edge leaves (HORZ at the bottom, VERT at the right), each for 1 bit, and with neither rows nor columns it
    must split (0 bits)."""

    def __init__(self, r, lam, avg_part, shape_mask=0x3F1FFF, part8=0xF, partN=0x3FF, max_part=64, split_bias=1.0):
        self.FW, self.FH = r["MiCols"] * 4, r["MiRows"] * 4
        self.br = r["blockres"]
        self.lam, self.ap = float(lam), avg_part.astype(np.float64)
        self.mask, self.p8, self.pN, self.maxp, self.bias = shape_mask, part8, partN, max_part, split_bias

    def block_J(self, w, h, x, y):
        bs = BS.get((w, h))
        if bs is None or not (self.mask >> bs) & 1:
            return None
        nbx = (self.FW + w - 1) // w
        a = self.br[bs]
        j = float(a["J"][(y // h) * nbx + x // w])
        return None if j >= 1e37 else j

    def trees(self, n, x, y, sub=None):
        """Every legal tree of the node: list of (cost, leaves); sub(n, x, y) gives the trees of a sub-node (default:
        all of them)."""
        sub = sub or self.trees
        FW, FH = self.FW, self.FH
        if x >= FW or y >= FH:
            return [(0.0, [])]
        h = n // 2
        has_rows, has_cols = y + h < FH, x + h < FW
        bsl = {8: 1, 16: 2, 32: 3, 64: 4}[n]
        pm = self.p8 if n == 8 else self.pN
        out = []
        if has_rows and has_cols:
            cands = [(SPLIT, self.ap[bsl][SPLIT])]
            if n <= self.maxp:
                cands += [(p, self.ap[bsl][p]) for p in (NONE, HORZ, VERT, HORZ_A, HORZ_B, VERT_A, VERT_B, HORZ_4, VERT_4)
                          if (pm >> p) & 1 and (n > 8 or p <= SPLIT)]
        elif has_cols:
            cands = [(SPLIT, 1.0)] + ([(HORZ, 1.0)] if n <= self.maxp else [])
        elif has_rows:
            cands = [(SPLIT, 1.0)] + ([(VERT, 1.0)] if n <= self.maxp else [])
        else:
            cands = [(SPLIT, 0.0)]
        for p, bits in cands:
            parts = partition_blocks(p, n, x, y)
            combos = [(self.lam * bits, [])]
            ok = True
            for b in parts:
                if b[0] == "node":
                    st = sub(b[1], b[2], b[3])
                    combos = [(c + c2, l + l2) for c, l in combos for c2, l2 in st]
                    continue
                w, hh, bx, by = b
                if bx >= FW or by >= FH:
                    continue
                j = self.block_J(w, hh, bx, by)
                if j is None:
                    ok = False
                    break
                if p == NONE:
                    j *= self.bias
                combos = [(c + j, l + [(BS[(w, hh)], bx, by)]) for c, l in combos]
            if ok:
                out += combos
        return out

    def best(self, n, x, y, full_below=16):
        """(cost, leaves) of the minimum tree: nodes up to full_below by enumerating every tree, larger ones by
        trying every partition over the optimal sub-nodes."""
        if n <= full_below:
            return min(self.trees(n, x, y), key=lambda z: z[0])
        return min(self.trees(n, x, y, sub=lambda nn, xx, yy: [self.best(nn, xx, yy, full_below)]), key=lambda z: z[0])


def _leaves_of(r, sb, sbx, sby):
    n = int(r["leaf_count"][sb])
    lv = r["leaves"][sb][:n]
    return [(int(b["bsize"]), sbx + 4 * int(b["x4"]), sby + 4 * int(b["y4"])) for b in lv]


@pytest.mark.parametrize("W,H,seed,q,kw", [
    (64, 64, 11, 120, {}), (136, 72, 12, 60, {}), (40, 24, 13, 160, {}), (72, 136, 14, 200, {}),
    (96, 64, 15, 100, {"partMaskN": 0xF0}), (64, 64, 16, 120, {"maxPart": 32}), (88, 56, 17, 90, {"splitBias": 1.3}),
    (64, 64, 18, 120, {"shapeMask": 0x3F1FFF & ~((1 << 9) | (1 << 6))}), (24, 16, 19, 120, {})])
def test_dp_equals_brute_force(oracle, W, H, seed, q, kw):
    import av1gpu_amd as A
    planes = A.synth_image(2000 + seed, W, H)
    r = oracle.encode(planes, W, H, qindex=q, params=oracle.params(**kw), detail=True)
    _, lam, ap = txg_table(oracle)
    p = oracle.params(**kw)
    B = Brute(r, lam[0][q] * p.lambdaMul, ap, p.shapeMask, p.partMask8, p.partMaskN, p.maxPart, p.splitBias)
    sbc = (r["MiCols"] + 15) >> 4
    for sb in range(len(r["sbJ"])):
        sbx, sby = (sb % sbc) * 64, (sb // sbc) * 64
        cost, leaves = B.best(64, sbx, sby)
        assert abs(float(r["sbJ"][sb]) - cost) <= 1e-5 * abs(cost) + 1e-3, (sb, r["sbJ"][sb], cost)
        got = _leaves_of(r, sb, sbx, sby)
        if got != leaves:   # only an exact tie in float64 terms may pick another tree of (nearly) the same cost
            tree_cost = sum(B.block_J(*{v: k for k, v in BS.items()}[b], x, y) for b, x, y in got)
            assert abs(tree_cost - sum(B.block_J(*{v: k for k, v in BS.items()}[b], x, y) for b, x, y in leaves)) <= 1e-3 * abs(cost) + 1e-2


def test_dp_full_enumeration_small_frames(oracle):
    """Every partition tree of 32x32 and 32x24 frames with 8x8 the smallest shape (so the trees stay countable):
    the oracle's superblock cost is the minimum over all of them."""
    import av1gpu_amd as A
    mask = sum(1 << b for b in (3, 4, 5, 6, 7, 8, 9, 18, 19))   # shapes with both sides >= 8
    for W, H, q in ((32, 32, 120), (32, 24, 60)):
        planes = A.synth_image(2100 + W + H, W, H)
        r = oracle.encode(planes, W, H, qindex=q, params=oracle.params(shapeMask=mask, partMask8=0x1), detail=True)
        _, lam, ap = txg_table(oracle)
        B = Brute(r, lam[0][q], ap, mask, 0x1, 0x3FF)
        trees = B.trees(64, 0, 0)
        assert len(trees) > 200
        best = min(z[0] for z in trees)
        assert abs(float(r["sbJ"][0]) - best) <= 1e-5 * best + 1e-3
