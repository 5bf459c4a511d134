"""Known answer for the oracle's coefficient rate model (the R of every RD cost), taken from the reference's own
writer: for one coded tx block, R must equal the ideal cost of the symbols the reference's bitstream writer codes for
it -- -log2 of each symbol's probability under the fresh default CDF at the cdf id the writer chose
(av1core/src/encoder.cpp:197-214 TeeSink, 1580-1584; cdfctx.h av1g_cdfs_init), plus one bit per equiprobable bit (sign,
eob extra and Golomb bits) -- for every tx size and intra tx type, luma and chroma, all four coefficient-CDF
q-contexts and random levels (sparse and dense, with levels past the Golomb threshold).  The writer derives the
contexts (all_zero, eob, coeff_base, coeff_br, dc_sign) itself, so a wrong context, table index or symbol set in the
oracle's rate model changes the sum.  Also the quantiser's closed forms (dead zone, rounding, clamp, eob) and its
distortion against float64, and the forward transform against the exact inverse (flips, scale, 64-point truncation)."""
import numpy as np
import pytest

import refcheck

# AV1 spec tables (transform and block sizes, intra tx sets; spec 6.10.19, 5.11.47, 7.13.3, Annex tables)
TXW = [4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64]
TXH = [4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16]
SIZES = {4: 0, 8: 1, 16: 2, 32: 3, 64: 4}
BSIZE = {(1, 1): 0, (1, 2): 1, (2, 1): 2, (2, 2): 3, (2, 4): 4, (4, 2): 5, (4, 4): 6, (4, 8): 7, (8, 4): 8, (8, 8): 9,
         (8, 16): 10, (16, 8): 11, (16, 16): 12, (1, 4): 16, (4, 1): 17, (2, 8): 18, (8, 2): 19, (4, 16): 20, (16, 4): 21}
IDTX, DCT_DCT, ADST_DCT, DCT_ADST, ADST_ADST, V_DCT, H_DCT = 9, 0, 1, 2, 3, 10, 11
SET1 = [IDTX, DCT_DCT, V_DCT, H_DCT, ADST_ADST, ADST_DCT, DCT_ADST]
SET2 = [IDTX, DCT_DCT, ADST_ADST, ADST_DCT, DCT_ADST]


def intra_set(t):
    sqr_up, sqr = max(TXW[t], TXH[t]), min(TXW[t], TXH[t])
    if sqr_up >= 32:
        return [DCT_DCT]
    return SET2 if sqr == 16 else SET1


def tx_size_ctx(t):
    return (SIZES[min(TXW[t], TXH[t])] + SIZES[max(TXW[t], TXH[t])] + 1) >> 1


def blocks_with_first(w4, h4, n4=16, x4=0, y4=0):
    """Blocks (mi_row, mi_col, bsize) in decode order of a 64x64 superblock whose first block is w4 x h4 at (0, 0): the
    partition of the smallest square node that yields that shape first (NONE / HORZ / VERT / HORZ_4 / VERT_4), SPLIT
    above it, and NONE in every other quadrant."""
    out = []
    if (w4, h4) == (n4, n4):
        return [(y4, x4, BSIZE[(n4, n4)])]
    for k, (a, b) in enumerate(((n4, n4 // 2), (n4 // 2, n4), (n4, n4 // 4), (n4 // 4, n4))):
        if (w4, h4) == (a, b) and BSIZE.get((a, b)) is not None and (n4 > 1):
            if k in (0, 2):
                return [(y4 + i * b, x4, BSIZE[(a, b)]) for i in range(n4 // b)]
            return [(y4, x4 + i * a, BSIZE[(a, b)]) for i in range(n4 // a)]
    h = n4 // 2
    if h == 0:
        raise ValueError("no partition")
    if n4 == 2:   # the 8x8 node split: four 4x4 blocks
        return [(y4 + dy, x4 + dx, 0) for dy in (0, 1) for dx in (0, 1)]
    out += blocks_with_first(w4, h4, h, x4, y4)
    for dy, dx in ((0, h), (h, 0), (h, h)):
        out.append((y4 + dy, x4 + dx, BSIZE[(h, h)]))
    return out


def make_blocks(spec, y_mode=0):
    b = np.zeros(len(spec), refcheck.BLOCK_DTYPE)
    for i, (r, c, bs) in enumerate(spec):
        b[i]["mi_row"], b[i]["mi_col"], b[i]["bsize"], b[i]["skip"] = r, c, bs, 1
    b[0]["skip"] = 0
    b[0]["y_mode"] = y_mode
    return b


def random_levels(rng, n, density, big):
    q = np.zeros(n, np.int16)
    k = max(1, int(round(density * n)))
    pos = rng.choice(n, k, replace=False)
    mags = rng.choice([1, 1, 1, 2, 2, 3, 4, 5, 8, 12], k)
    if big:
        sel = rng.random(k) < 0.15
        mags[sel] = rng.integers(15, 60, sel.sum())
    q[pos] = mags * rng.choice([-1, 1], k)
    return q


class SymbolCoster:
    def __init__(self, ref, qindex):
        self.cdf = ref.default_cdfs(qindex).astype(np.float64)
        self.fields = ref.coeff_cdf_fields()

    def field_of(self, off):
        for f, (o, n) in self.fields.items():
            if o <= off < o + n:
                return f, off - o
        return None, None

    def segment(self, syms, k):
        """The symbols of the k-th coded tx block (from its all_zero symbol to the next all_zero or to a symbol of
        another syntax element) and their ideal cost in bits under the default CDFs."""
        starts = [i for i, w in enumerate(syms) if (w >> 24) == 0 and self.field_of(int(w & 0xFFFF))[0] == "txb_skip"]
        i = starts[k]
        seg = [int(syms[i])]
        for w in syms[i + 1:]:
            w = int(w)
            kind = w >> 24
            f = self.field_of(w & 0xFFFF)[0] if kind == 0 else None
            if kind == 1 or (kind == 0 and f is not None and f != "txb_skip"):
                seg.append(w)
            else:
                break
        bits = 0.0
        for w in seg:
            kind, s, off = w >> 24, (w >> 16) & 15, w & 0xFFFF
            if kind == 1:
                bits += 1.0
            else:
                lo = self.cdf[off + s - 1] if s else 0.0
                p = max((self.cdf[off + s] - lo) / 32768.0, 1.0 / 32768.0)
                bits -= np.log2(p)
        return seg, bits


def _ctx_of(coster, seg):
    f, rel = coster.field_of(seg[0] & 0xFFFF)
    assert f == "txb_skip"
    idx = rel // 3
    return idx // 13, idx % 13


CASES = [(t, tt) for t in range(19) for tt in intra_set(t)]


@pytest.mark.parametrize("t,tx_type", CASES)
def test_rate_equals_writer_symbols_luma(oracle, ref, t, tx_type):
    rng = np.random.default_rng(1000 + 31 * t + tx_type)
    w4, h4 = TXW[t] // 4, TXH[t] // 4
    spec = blocks_with_first(w4, h4)
    tw, th = min(TXW[t], 32), min(TXH[t], 32)
    for qindex, bd in ((20, 8), (60, 10), (120, 8), (255, 8)):
        coster = SymbolCoster(ref, qindex)
        for density, big in ((0.0, False), (0.05, False), (0.3, True), (1.0, True)):
            y_mode = int(rng.integers(0, 13))
            blocks = make_blocks(spec, y_mode)
            bt, txb, nco = ref.layout(blocks, 64, 64, qindex, mono=1, bd=bd)
            assert txb[0]["tx_size"] == t and txb[0]["plane"] == 0
            coeffs = np.zeros(nco, np.int16)
            q = random_levels(rng, tw * th, density, big) if density > 0 else np.zeros(tw * th, np.int16)
            coeffs[txb[0]["coeff_off"]:txb[0]["coeff_off"] + tw * th] = q
            txb = txb.copy()
            txb[0]["tx_type"] = tx_type
            syms = ref.symbols(dict(blocks=blocks, block_txb=bt, txb=txb, coeffs=coeffs), 64, 64, qindex, mono=1, bd=bd)
            seg, bits = coster.segment(syms, 0)
            szctx, skip_ctx = _ctx_of(coster, seg)
            assert szctx == tx_size_ctx(t) and skip_ctx == 0   # a tx block covering its block, no neighbours
            R, eob = oracle.rate_levels(q, t, tx_type, 0, qindex, skip_ctx, 0, y_mode)
            assert (eob == 0) == (not q.any())
            assert abs(R - bits) <= 1e-3 + 1e-5 * bits, (qindex, density, R, bits)


CHROMA = [(w, h) for w, h in BSIZE if w >= 2 and h >= 2]


@pytest.mark.parametrize("w4,h4", CHROMA)
@pytest.mark.parametrize("ssx", [1, 0])
def test_rate_equals_writer_symbols_chroma(oracle, ref, w4, h4, ssx):
    rng = np.random.default_rng(7 * w4 + 131 * h4 + ssx)
    spec = blocks_with_first(w4, h4)
    for qindex in (40, 100, 200):
        coster = SymbolCoster(ref, qindex)
        for uv_mode, density, big in ((0, 0.2, False), (1, 0.5, True), (2, 1.0, True), (9, 0.1, True)):
            blocks = make_blocks(spec)
            blocks[0]["uv_mode"] = uv_mode
            bt, txb, nco = ref.layout(blocks, 64, 64, qindex, ssx=ssx)
            k = int(np.nonzero(txb["plane"][:bt[1]] == 1)[0][0])      # block 0's first U tx block
            t = int(txb[k]["tx_size"])
            tw, th = min(TXW[t], 32), min(TXH[t], 32)
            coeffs = np.zeros(nco, np.int16)
            q = random_levels(rng, tw * th, density, big)
            coeffs[txb[k]["coeff_off"]:txb[k]["coeff_off"] + tw * th] = q
            rec = dict(blocks=blocks, block_txb=bt, txb=txb, coeffs=coeffs)
            syms = ref.symbols(rec, 64, 64, qindex, ssx=ssx)
            # the chroma tx type is derived from the uv mode by the writer: read it back from an encode of the same
            # blocks without coefficients (no range limits to meet)
            _, _, info = ref.encode(dict(rec, coeffs=np.zeros_like(coeffs)), 64, 64, qindex, ssx=ssx, layout_check=False)
            tx_type = int(info["tx_type"][k])
            seg, bits = coster.segment(syms, k)
            szctx, skip_ctx = _ctx_of(coster, seg)
            assert szctx == tx_size_ctx(t) and skip_ctx >= 7   # chroma all_zero contexts start at 7 (spec get_tx_block_ctx)
            R, eob = oracle.rate_levels(q, t, tx_type, 1, qindex, skip_ctx, 0, 0)
            assert eob > 0
            assert abs(R - bits) <= 1e-3 + 1e-5 * bits, (qindex, uv_mode, R, bits)


def test_rate_model_detects_a_wrong_context(oracle, ref):
    """The same comparison fails when the oracle is given a context other than the writer's (the check has teeth)."""
    rng = np.random.default_rng(5)
    t = 1
    spec = blocks_with_first(2, 2)
    coster = SymbolCoster(ref, 120)
    blocks = make_blocks(spec)
    bt, txb, nco = ref.layout(blocks, 64, 64, 120, mono=1)
    q = random_levels(rng, 64, 0.5, True)
    coeffs = np.zeros(nco, np.int16)
    coeffs[:64] = q
    syms = ref.symbols(dict(blocks=blocks, block_txb=bt, txb=txb, coeffs=coeffs), 64, 64, 120, mono=1)
    _, bits = coster.segment(syms, 0)
    R0, _ = oracle.rate_levels(q, t, DCT_DCT, 0, 120, 0, 0, 0)
    R1, _ = oracle.rate_levels(q, t, DCT_DCT, 0, 120, 1, 0, 0)
    R2, _ = oracle.rate_levels(q, t, DCT_DCT, 0, 120, 0, 1, 0)
    assert abs(R0 - bits) <= 1e-3 + 1e-5 * bits
    assert abs(R1 - bits) > 1e-2 and (abs(R2 - bits) > 1e-2 or q[0] == 0)
