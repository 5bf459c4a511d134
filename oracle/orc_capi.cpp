// Oracle (test infrastructure only): C entry points of the CPU restatement for tests/ (ctypes) and bench.py's CPU
// baseline.  Nothing in the product links or loads this library.
#include "orc_search.h"
#include <math.h>
#include <stdlib.h>
#include <vector>

using namespace orc;

extern "C" {

struct OrcParams {   // the plain-search knobs (GpParams fields of the same names, gp.h:29-68)
    float lambdaMul, satdLambda, rnd, zbin, rndC, zbinC;
    int K, Krecheck, txSearch, txSearchCL, txAllCL, deltas, uvDeltas, cfl, uvK, shapeMask, partMask8, partMaskN, maxPart;
    float uvOverheadBits, splitBias, lambdaCMul;
};

void orc_default_params(OrcParams *p) {
    Params d;
    p->lambdaMul = d.lambdaMul; p->satdLambda = d.satdLambda; p->rnd = d.rnd; p->zbin = d.zbin; p->rndC = d.rndC; p->zbinC = d.zbinC;
    p->K = d.K; p->Krecheck = d.Krecheck; p->txSearch = d.txSearch; p->txSearchCL = d.txSearchCL; p->txAllCL = d.txAllCL;
    p->deltas = d.deltas; p->uvDeltas = d.uvDeltas; p->cfl = d.cfl; p->uvK = d.uvK; p->shapeMask = d.shapeMask;
    p->partMask8 = d.partMask8; p->partMaskN = d.partMaskN; p->maxPart = d.maxPart;
    p->uvOverheadBits = d.uvOverheadBits; p->splitBias = d.splitBias; p->lambdaCMul = d.lambdaCMul;
}
static Params to_params(const OrcParams *p) {
    Params d;
    if (!p) return d;
    d.lambdaMul = p->lambdaMul; d.satdLambda = p->satdLambda; d.rnd = p->rnd; d.zbin = p->zbin; d.rndC = p->rndC; d.zbinC = p->zbinC;
    d.K = p->K; d.Krecheck = p->Krecheck; d.txSearch = p->txSearch; d.txSearchCL = p->txSearchCL; d.txAllCL = p->txAllCL;
    d.deltas = p->deltas; d.uvDeltas = p->uvDeltas; d.cfl = p->cfl; d.uvK = p->uvK; d.shapeMask = p->shapeMask;
    d.partMask8 = p->partMask8; d.partMaskN = p->partMaskN; d.maxPart = p->maxPart;
    d.uvOverheadBits = p->uvOverheadBits; d.splitBias = p->splitBias; d.lambdaCMul = p->lambdaCMul;
    return d;
}

// Encode one frame.  planes: 8-bit (bd 8, uint8) or 16-bit (bd 10, uint16) samples, visible size (chroma
// (W + ssx) >> ssx wide), strides in samples.  Returns a handle (free with orc_free).
void *orc_encode(const void *y, const void *u, const void *v, const int *strides, int W, int H, int ssx, int mono, int bd,
                 int qindex, int edgeFilter, const OrcParams *params, int threads);
// The same with the analysis hooks: ext[s] (NULL or the external candidates of shape s, extTopK per block of the frame's
// grid in raster order) and analyse[s] (NULL or one byte per block of the grid: 1 = analysed).
void *orc_encode_hooks(const void *y, const void *u, const void *v, const int *strides, int W, int H, int ssx, int mono, int bd,
                       int qindex, int edgeFilter, const OrcParams *params, int threads, const void *const *ext, int extTopK,
                       const uint8_t *const *analyse);
void *orc_encode(const void *y, const void *u, const void *v, const int *strides, int W, int H, int ssx, int mono, int bd,
                 int qindex, int edgeFilter, const OrcParams *params, int threads) {
    return orc_encode_hooks(y, u, v, strides, W, H, ssx, mono, bd, qindex, edgeFilter, params, threads, nullptr, 0, nullptr);
}
void *orc_encode_hooks(const void *y, const void *u, const void *v, const int *strides, int W, int H, int ssx, int mono, int bd,
                       int qindex, int edgeFilter, const OrcParams *params, int threads, const void *const *ext, int extTopK,
                       const uint8_t *const *analyse) {
    Frame *F = new Frame();
    for (int s = 0; s < 22; s++) {
        F->ext[s] = ext ? (const ExtCand *)ext[s] : nullptr;
        F->analyse[s] = analyse ? analyse[s] : nullptr;
    }
    F->extTopK = extTopK;
    const void *pl[3] = {y, u, v};
    std::vector<uint16_t> tmp[3];
    const uint16_t *src[3];
    for (int p = 0; p < 3; p++) {
        const int sx = p ? ssx : 0, w = (W + sx) >> sx, h = (H + sx) >> sx;
        if (p && mono) { tmp[p].assign(1, 0); src[p] = tmp[p].data(); continue; }
        tmp[p].resize((size_t)strides[p] * h);
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++)
                tmp[p][(size_t)i * strides[p] + j] = bd == 8 ? ((const uint8_t *)pl[p])[(size_t)i * strides[p] + j]
                                                             : ((const uint16_t *)pl[p])[(size_t)i * strides[p] + j];
        src[p] = tmp[p].data();
    }
    frame_setup(*F, W, H, ssx, mono, bd, qindex, edgeFilter, src, strides);
    encode_frame(*F, to_params(params), threads);
    return F;
}
void orc_free(void *h) { delete (Frame *)h; }

// counts: n_blocks, n_txb, n_coeffs, MiCols, MiRows, aw, ah, tileCols, tileRows, stats[0..3], forward-transform FLOPs
// of the analysis and of the closed loop
void orc_counts(void *h, int64_t *out) {
    const Frame &F = *(Frame *)h;
    out[0] = (int64_t)F.blocks.size(); out[1] = (int64_t)F.txb.size(); out[2] = (int64_t)F.coeffs.size();
    out[3] = F.MiCols; out[4] = F.MiRows; out[5] = F.aw; out[6] = F.ah; out[7] = F.tileCols; out[8] = F.tileRows;
    for (int i = 0; i < 4; i++) out[9 + i] = F.stats[i];
    out[13] = (int64_t)F.flops[0].load(); out[14] = (int64_t)F.flops[1].load();
}
// record arrays (16-byte blocks / txbs, block_txb[n_blocks + 1], int16 coefficients)
void orc_record(void *h, void *blocks, uint32_t *blockTxb, void *txb, int16_t *coeffs) {
    const Frame &F = *(Frame *)h;
    memcpy(blocks, F.blocks.data(), F.blocks.size() * sizeof(RecBlock));
    memcpy(blockTxb, F.blockTxb.data(), F.blockTxb.size() * 4);
    memcpy(txb, F.txb.data(), F.txb.size() * sizeof(RecTxb));
    memcpy(coeffs, F.coeffs.data(), F.coeffs.size() * 2);
}
// padded reconstruction plane p (aw >> sx x ah >> sy samples, uint16)
void orc_recon(void *h, int p, uint16_t *out) {
    const Frame &F = *(Frame *)h;
    memcpy(out, F.rec[p].data(), F.rec[p].size() * 2);
}
// analysis records of one shape: nbx * nby entries of {J, cand[6], ncand, tx} (12 bytes)
int orc_blockres(void *h, int shape, void *out) {
    const Frame &F = *(Frame *)h;
    if (shape < 0 || shape >= 22) return 0;
    memcpy(out, F.blk[shape].data(), F.blk[shape].size() * sizeof(BlockRes));
    return (int)F.blk[shape].size();
}
// leaves of every SB: per SB the count, then 256 x {bsize, x4, y4, pad}
void orc_leaves(void *h, uint16_t *counts, void *leaves) {
    const Frame &F = *(Frame *)h;
    for (size_t s = 0; s < F.leaves.size(); s++) {
        counts[s] = (uint16_t)F.leaves[s].size();
        memcpy((Leaf *)leaves + s * 256, F.leaves[s].data(), F.leaves[s].size() * sizeof(Leaf));
    }
}
void orc_sbj(void *h, float *out) {
    const Frame &F = *(Frame *)h;
    memcpy(out, F.sbJ.data(), F.sbJ.size() * 4);
}

// ---- primitives for the per-function parity tests ----
// forward matrix of kind (0 DCT, 1 ADST, 2 identity) and log2 size; returns n * n or 0
int orc_fwd_matrix(int kind, int ln, float *out) {
    const Tables &T = tables();
    if (kind < 0 || kind > 2 || ln < 2 || ln > 6 || T.fwd[kind][ln].empty()) return 0;
    memcpy(out, T.fwd[kind][ln].data(), T.fwd[kind][ln].size() * 4);
    return (int)T.fwd[kind][ln].size();
}
// txg[19][16], txScale[19], avgYMode[13], avgPart[5][10], avgSkip0, lambda[2][256], lamS[2][256]
void orc_misc_tables(float *txg, float *txScale, float *avgY, float *avgPart, float *avgSkip0, float *lambda, float *lamS) {
    const Tables &T = tables();
    memcpy(txg, T.txg, sizeof T.txg); memcpy(txScale, T.txScale, sizeof T.txScale); memcpy(avgY, T.avgYMode, sizeof T.avgYMode);
    memcpy(avgPart, T.avgPart, sizeof T.avgPart); *avgSkip0 = T.avgSkip0;
    memcpy(lambda, T.lambda, sizeof T.lambda); memcpy(lamS, T.lamS, sizeof T.lamS);
}
int orc_cost_table(int qctx, float *out) {
    memcpy(out, &tables().cost[qctx], sizeof(CostTab));
    return (int)(sizeof(CostTab) / 4);
}
// evaluation of one tx block; out: D, R, eob (as float); q: levels; coef: the forward transform outputs
void orc_eval_txb(const int16_t *res, int txSz, int txType, int plane, int qindex, int bd, int skipCtx, int dcCtx, int intraDir,
                  float rnd, float zbin, float *out, int16_t *q, float *coef) {
    EvalIn e;
    e.txSz = txSz; e.txType = txType; e.plane = plane; e.dcQ = dc_q(qindex, bd); e.acQ = ac_q(qindex, bd); e.rnd = rnd; e.zbin = zbin;
    e.skipCtx = skipCtx; e.dcCtx = dcCtx; e.intraDir = intraDir; e.bd = bd;
    e.C = &tables().cost[qindex <= 20 ? 0 : qindex <= 60 ? 1 : qindex <= 120 ? 2 : 3];
    const int W = txw(txSz), H = txh(txSz);
    e.resEnergy = (W == 64 || H == 64) ? res_energy(res, W * H) : 0.f;
    const EvalOut o = eval_txb(res, e, q, coef);
    out[0] = o.D; out[1] = o.R; out[2] = (float)o.eob;
}
// coefficient rate (bits) of levels q (row-major tw x th) of one tx block with the given contexts; *eob: its eob
float orc_rate_levels(const int16_t *q, int txSz, int txType, int plane, int qindex, int skipCtx, int dcCtx, int intraDir, int *eob) {
    EvalIn e;
    e.txSz = txSz; e.txType = txType; e.plane = plane; e.dcQ = e.acQ = 1; e.rnd = e.zbin = 0.f;
    e.skipCtx = skipCtx; e.dcCtx = dcCtx; e.intraDir = intraDir; e.bd = 8; e.resEnergy = 0.f;
    e.C = &tables().cost[qindex <= 20 ? 0 : qindex <= 60 ? 1 : qindex <= 120 ? 2 : 3];
    return rate_of_levels(q, e, eob);
}
// dequantize + inverse transform (no repair); res: h x w; returns the conformance flags
uint32_t orc_dequant_inverse(const int16_t *q, int txSz, int txType, int dcQ, int acQ, int bd, int32_t *res) {
    int32_t dq[1024];
    const int n = tx_coeffs(txSz), den = dq_denom(txSz);
    for (int i = 0; i < n; i++) dq[i] = q[i] ? dequant_coef(q[i], i == 0 ? dcQ : acQ, den, bd) : 0;
    uint32_t err = 0;
    inv_txfm2d(dq, txSz, txType, res, &err, bd);
    return err;
}
// intra prediction of a tx block from explicit edges (above/left: EDGE_LEN arrays, index EDGE_OFF is sample 0)
void orc_predict(int mode, int delta, int edgeFilter, int filterType, int haveL, int haveA, int txSz, int x, int y, int maxX,
                 int maxY, const uint16_t *above, const uint16_t *left, int bd, int *pred) {
    const int lw = av1t::Tx_Width_Log2[txSz], lh = av1t::Tx_Height_Log2[txSz];
    const int dcv = dc_value(above + EDGE_OFF, left + EDGE_OFF, haveL, haveA, lw, lh, bd);
    predict_block(mode, delta, edgeFilter, filterType, haveL, haveA, lw, lh, x, y, maxX, maxY, above + EDGE_OFF, left + EDGE_OFF, dcv,
                  pred, bd);
}
void orc_build_edges(const uint16_t *plane, int stride, int x, int y, int txSz, int haveL, int haveA, int haveAR, int haveBL, int maxX,
                     int maxY, int bd, uint16_t *above, uint16_t *left) {
    build_edges(plane, stride, x, y, av1t::Tx_Width_Log2[txSz], av1t::Tx_Height_Log2[txSz], haveL, haveA, haveAR, haveBL, maxX, maxY,
                above + EDGE_OFF, left + EDGE_OFF, bd);
}
int orc_tile_layout(int W, int H, int *cols, int *rows, int *colStart, int *rowStart) {
    return tile_layout(W, H, cols, rows, colStart, rowStart);
}

}  // extern "C"

extern "C" {
// one 1-D inverse transform (kind 0 DCT / 1 ADST / 2 identity, log2 size n, clamp range r); returns the flags
uint32_t orc_inv_1d(int32_t *T, int kind, int n, int r) {
    uint32_t err = 0;
    inv_1d(T, kind, n, r, &err);
    return err;
}
}
