// Closed-loop decisions and reconstruction: one 256-thread workgroup (8 groups of 32 lanes) per 64x64 superblock,
// launched per wavefront step (SBs whose left, above and above-right neighbours are complete).  Inside the SB the
// leaves run in decode order exactly as the decoder sees them: BlockDecoded state, real reconstructed edges, exact
// coefficient contexts from the 4x4 maps.  Reference: encoder/src/k_recon.cuh:1692-1808 (kernel), 822-1213 (luma),
// 1220-1686 (chroma), 139-197 (inverse + add), 203-238 (contexts), 685-730 (records).
//
// Per block:
//   luma    every re-checked analysis candidate x every intra tx type of the set (tx blocks up to 16x16; DCT_DCT only
//           above), RD tasks spread over the groups, the first minimum of J = fmaf(lambda, R + mode bits, D) wins;
//   chroma  CfL alpha (least squares from the reconstructed luma, then +-1 by exact SSE), SATD screening of the 13 UV
//           modes and the deltas around the best directional one, top uvK (+ CfL) RD tasks for U and V, pick;
//   final   the winner re-quantized (the same levels), prediction, exact dequant + inverse transform + add with the
//           conformance repair, tx records, context maps, BlockDecoded, block record, 4x4 maps.
#include "dev_eval.h"

namespace av1m {

constexpr int RC_NT = 256, RC_NG = 8;
// transform teams for tx blocks with a side of 32 or 64 (groups 0 .. nTeam - 1): RC_NTEAM of them inside RcShared keep
// the workgroup's LDS under half a compute unit's, so two workgroups share a CU; the one-workgroup-per-CU instance adds
// RC_NTEAM more after RcShared (every group a team)
constexpr int RC_NTEAM = 4;

constexpr int RC_SMALL_RES = 16;
// levels of every task of a tx block up to 16x16, packed by the block's size (task t at t * W * H): at most 6 candidates
// x 7 tx types of 128 coefficients, 6 x 5 of 256 (16x16: intra set 2), or 2 planes x 7 chroma candidates of 256
constexpr int RC_LEV_TOTAL = 7680;
// per-group scratch of the evaluations of transform blocks up to 16x16
struct RcGroup {
    float tmp[16 * 17];
    uint16_t ea[EDGE_LEN], el[EDGE_LEN];
};
// blocks up to 16x16: the residual of each candidate (luma) or task (chroma) and the levels of every task, kept until
// the decision so the winner is neither predicted nor quantized again
struct RcSmall {
    RcGroup g[RC_NG];
    int16_t res[RC_SMALL_RES][256];
    int16_t lev[RC_LEV_TOTAL];
};
// scratch of one group evaluating a transform block with a side of 32 or 64 (DCT_DCT only)
struct RcTeam {
    float tmp[32 * 65];      // th x (W + 1): column-pass outputs, then the coefficients
    int16_t q[1024];
    uint16_t ea[EDGE_LEN], el[EDGE_LEN];
};
// the tables the small-block evaluations and the decisions read, staged once per workgroup
struct RcTabs {
    float fwdT[3 * 336];     // DCT, ADST, identity for n = 4, 8, 16, transposed (offsets fwdOff)
    float fwd[3 * 336];      // the same, row-major (the identity kind reads its diagonal)
    int fwdOff[3][7];
    float txScale[19];
    float txg[19][16];
    uint8_t cbOff[19][25];         // Coeff_Base_Ctx_Offset
    const uint16_t *iscan;         // the inverse default scans (device tables; read through the caches)
    const CostTab *cost;           // the cost tables of the image's quantizer context (device tables)
};
using RcEvalIn = EvalInT<const CostTab *>;
template <int TSZ>
__device__ __forceinline__ const uint16_t *dscan_ptr(const lds<RcTabs> &) { return scan_of(TSZ, DCT_DCT); }
template <int TSZ>
__device__ __forceinline__ const lds<uint8_t> *cboff_ptr(const lds<RcTabs> &T) { return T.cbOff[TSZ]; }
template <int TSZ>
__device__ __forceinline__ const uint16_t *iscan_ptr(const lds<RcTabs> &T) { return T.iscan + dscan_off(av1t::Adjusted_Tx_Size[TSZ]); }

union RcScr {
    RcSmall s;
    RcTeam t[RC_NTEAM];
};
struct RcShared {
    RcScr u;
    RcTabs tab;
    uint16_t srcb[64 * 64];                  // source block of the current leaf: luma W x H, or U at 0 and V at 2048
    int16_t inv[64 * 64];                    // inverse transform residual (the row pass clamps to 16 bits; the column
                                             // outputs of a block passing the range checks are within 12)
    uint16_t ea[3][EDGE_LEN], el[3][EDGE_LEN];   // unfiltered edges of the current tx block per plane
    int dcv[3], ft[3], hvL[3], hvA[3], px[3], py[3], mx[3], my[3];
    int skipCtx[3], dcCtx[3];
    int gup[RC_NG][2];                       // upsampling flags of each group's prepared edges
    uint8_t bdm[3][18][18];                  // BlockDecoded[plane][y + 1][x + 1]
    int taskC[64], taskT[64], taskP[64], taskE[64], taskK[64];   // candidate, tx type, plane, eob, residual slot
    int candC[8], nCand;                     // luma: the re-checked candidates
    float taskJ[64];
    int nTask;
    float gBestJ[RC_NG];
    int gBestT[RC_NG];
    int win, winEob[3];
    int16_t cflL[32 * 32];                   // CfL luma (at most 8 x 1023)
    int cflAvg, cflAlpha[2];
    unsigned long long acc64[8];
    int satdI[32], uvc[32];
    float uvJ[32];                           // screening costs of the UV candidates
    int nTxb, nCoef, errFlag, nzFlag, eobOut, culOut;
    int nTeam;                               // transform teams: RC_NTEAM, or 2 * RC_NTEAM with the extra ones after RcShared
    int flopsOn;                             // count the forward-transform work (Batch::flops set)
    unsigned long long flops;                // of the superblock so far
    int am, lm, mode, delta, uvMode, uvDelta, nuv, bsize;
    uint32_t qItem;                          // persistent schedule: the superblock taken
};

// scratch of transform team g: the first RC_NTEAM share the union with the small-block scratch, the others follow RcShared
__device__ __forceinline__ lds<RcTeam> &team_scr(lds<RcShared> &S, int g) {
    return g < RC_NTEAM ? S.u.t[g] : ((lds<RcTeam> *)(lds<char> *)(&S + 1))[g - RC_NTEAM];
}
__device__ __forceinline__ const lds<RcTeam> &team_scr(const lds<RcShared> &S, int g) {
    return g < RC_NTEAM ? S.u.t[g] : ((const lds<RcTeam> *)(const lds<char> *)(&S + 1))[g - RC_NTEAM];
}

template <typename Px> __device__ __forceinline__ Px *rec_plane(const Img &I, int p) { return (Px *)I.rec[p]; }
template <typename Px> __device__ __forceinline__ const Px *src_plane(const Img &I, int p) { return (const Px *)I.src[p]; }

// get_filter_type (spec 7.11.2.8): 1 if the above or left block (chroma: the chroma-carrying one) uses a SMOOTH mode
__device__ int filter_type(const Img &I, int plane, int mi_row, int mi_col, int availU, int availL) {
    const uint8_t *m = plane ? I.mUV : I.mY;
    auto smooth = [](int v) { return v == SMOOTH_PRED || v == SMOOTH_V_PRED || v == SMOOTH_H_PRED; };
    int a = 0, l = 0;
    if (availU) {
        int r = mi_row - 1, c = mi_col;
        if (plane > 0 && I.ss) { if (!(mi_col & 1)) c++; if (mi_row & 1) r--; }
        a = smooth(m[(size_t)r * I.MiCols + c]);
    }
    if (availL) {
        int r = mi_row, c = mi_col - 1;
        if (plane > 0 && I.ss) { if (mi_col & 1) c--; if (!(mi_row & 1)) r++; }
        l = smooth(m[(size_t)r * I.MiCols + c]);
    }
    return a || l;
}

// coefficient contexts of a tx block from the context maps (spec get_tx_block ctx), by the 32 lanes of the group:
// lane k reads the k-th 4x4 column above and row to the left
__device__ void grp_txb_ctx(const Img &I, int plane, int txSz, int x4, int y4, int bsize, int tx0, int ty0, lds<int> *skipCtx,
                            lds<int> *dcCtx) {
    const int lane = glane(), ss = plane ? I.ss : 0;
    const int stride = I.MiCols >> ss, maxX4 = I.MiCols >> ss, maxY4 = I.MiRows >> ss;
    const uint8_t *L = I.cLvl[plane], *D = I.cDc[plane];
    const bool hasA = y4 > ((ty0 >> 2) >> ss), hasL = x4 > ((tx0 >> 2) >> ss);
    const int w = tx_w(txSz), h = tx_h(txSz), w4 = w >> 2, h4 = h >> 2;
    int top = 0, left = 0, topAny = 0, leftAny = 0, sgn = 0;
    if (hasA && lane < w4 && x4 + lane < maxX4) {
        const int l = L[(size_t)(y4 - 1) * stride + x4 + lane], d = D[(size_t)(y4 - 1) * stride + x4 + lane];
        top = l; topAny = (l | d) != 0; sgn += d == 2 ? 1 : d == 1 ? -1 : 0;
    }
    if (hasL && lane < h4 && y4 + lane < maxY4) {
        const int l = L[(size_t)(y4 + lane) * stride + x4 - 1], d = D[(size_t)(y4 + lane) * stride + x4 - 1];
        left = l; leftAny = (l | d) != 0; sgn += d == 2 ? 1 : d == 1 ? -1 : 0;
    }
    top = gmaxi(top); left = gmaxi(left); topAny = gmaxi(topAny); leftAny = gmaxi(leftAny); sgn = gsumi(sgn);
    if (lane == 0) {
        if (plane == 0) {
            const int bw = 4 * bw4(bsize), bh = 4 * bh4(bsize);
            *skipCtx = (bw == w && bh == h) ? 0 : (top == 0 && left == 0) ? 1 : (top == 0 || left == 0) ? 2 + (max(top, left) > 3)
                     : max(top, left) <= 3 ? 4 : min(top, left) <= 3 ? 5 : 6;
        } else {
            const int ub = uv_bsize(bsize, I.ss);
            *skipCtx = topAny + leftAny + 7 + ((16 * bw4(ub) * bh4(ub) > w * h) ? 3 : 0);
        }
        *dcCtx = sgn < 0 ? 1 : sgn > 0 ? 2 : 0;
    }
}

// base edges of the tx block of plane p at (px, py) from the reconstruction (group g), DC value and filter type
template <typename Px>
// fA / fL: the block's AvailU / AvailL for the edge filter type where they differ from the tx block's edges (a further
// 4:4:4 chroma tx block inside its block); -1: the same
__device__ void rc_edges(lds<RcShared> &S, const Img &I, int p, int px, int py, int lw, int lh, int haveL, int haveA, int haveAR, int haveBL,
                         int mi_row, int mi_col, int fA = -1, int fL = -1) {
    const int ss = p ? I.ss : 0;
    const int maxX = ((I.MiCols * 4) >> ss) - 1, maxY = ((I.MiRows * 4) >> ss) - 1;
    lds<uint16_t> *A = S.ea[p] + EDGE_OFF, *L = S.el[p] + EDGE_OFF;
    grp_build_edges(rec_plane<Px>(I, p), I.pitch[p], px, py, lw, lh, haveL, haveA, haveAR, haveBL, maxX, maxY, A, L, I.bd);
    const int dcv = grp_dc_value(A, L, haveL, haveA, lw, lh, I.bd);
    if (glane() == 0) {
        S.dcv[p] = dcv; S.hvL[p] = haveL; S.hvA[p] = haveA; S.px[p] = px; S.py[p] = py; S.mx[p] = maxX; S.my[p] = maxY;
        S.ft[p] = I.edgeFilter ? filter_type(I, p, mi_row, mi_col, fA < 0 ? haveA : fA, fL < 0 ? haveL : fL) : 0;
    }
}

// prediction of candidate c (255 = CfL) of plane p for a W x H tx block (W, H <= 16), by the group, into res = src -
// pred (raster order: lane l makes samples l, l + 32, ..., i.e. column l % W, rows l / W + k * 32 / W, as one run of
// pred_col); the group's edge buffers receive the prepared (filtered) edges of directional modes.  S.gup[gid] = flags.
template <int W, int H>
__device__ void grp_residual(lds<RcShared> &S, lds<uint16_t> *gea, lds<uint16_t> *gel, int p, int c, const lds<uint16_t> *src, int sp,
                             lds<int16_t> *res, int bd) {
    constexpr int lw = W == 4 ? 2 : W == 8 ? 3 : 4, lh = H == 4 ? 2 : H == 8 ? 3 : 4;
    constexpr int NPX = W * H, CH = NPX >= GW ? NPX / GW : 1, DI = GW / W;
    const int lane = glane();
    const int m = c == 255 ? DC_PRED : cand_mode(c), pa = pangle(m, c == 255 ? 0 : cand_delta(c));
    const lds<uint16_t> *EA = S.ea[p] + EDGE_OFF, *EL = S.el[p] + EDGE_OFF;
    int upA = 0, upL = 0;
    const bool filt = c != 255 && is_dir(m) && gea != nullptr;
    if (filt) {
        for (int i = lane; i < EDGE_LEN; i += GW) { gea[i] = S.ea[p][i]; gel[i] = S.el[p][i]; }
        gsync();
        grp_edge_prepare(pa, lw, lh, S.ft[p], S.hvL[p], S.hvA[p], S.px[p], S.py[p], S.mx[p], S.my[p], gea + EDGE_OFF, gel + EDGE_OFF, bd, upA,
                         upL);
        EA = gea + EDGE_OFF;
        EL = gel + EDGE_OFF;
    }
    const int dcv = S.dcv[p];
    if (lane < NPX) {
        const int i0 = lane >> lw, j = lane & (W - 1);
        int r[CH];
        if (c == 255) {
            const int alpha = S.cflAlpha[p - 1], avg = S.cflAvg;
#pragma unroll
            for (int k = 0; k < CH; k++) r[k] = clip3(0, (1 << bd) - 1, dcv + round2signed(alpha * (S.cflL[(i0 + k * DI) * W + j] - avg), 6));
        } else {
            pred_col(m, pa, dcv, i0, DI, j, lw, lh, EA, EL, upA, upL, r);
        }
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int i = i0 + k * DI;
            res[i * W + j] = (int16_t)((int)src[i * sp + j] - r[k]);
        }
    }
    if (lane == 0) { S.gup[gid()][0] = upA; S.gup[gid()][1] = upL; }
    gsync();
}

// one RD task (candidate c of plane p with tx type t) of a tx block up to 16x16 by the calling group: the residual
// res is ready; the levels go to lev
template <int W, int H>
__device__ EvalOut grp_task(lds<RcShared> &S, lds<RcGroup> &X, const Img &I, int p, int c, int t, const lds<int16_t> *res,
                            lds<int16_t> *lev, const SearchParams &P) {
    RcEvalIn e;
    e.txSz = tx_of(W, H); e.txType = t; e.plane = p;
    e.dcQ = dc_q(I.qindex, I.bd); e.acQ = ac_q(I.qindex, I.bd);
    e.skipCtx = S.skipCtx[p]; e.dcCtx = S.dcCtx[p]; e.intraDir = p == 0 ? cand_mode(c) : 0; e.bd = I.bd;
    e.rnd = p ? P.rndC : P.rnd; e.zbin = p ? P.zbinC : P.zbin;
    e.resEnergy = 0.f;
    e.C = S.tab.cost;
    return grp_eval<W, H>(res, X.tmp, lev, e, S.tab);
}
// the same task by half the group (blocks of up to 64 coefficients), with tmp its own part of the group's scratch
template <int W, int H>
__device__ EvalOut half_task(lds<RcShared> &S, lds<float> *tmp, const Img &I, int p, int c, int t, const lds<int16_t> *res,
                             lds<int16_t> *lev, const SearchParams &P) {
    RcEvalIn e;
    e.txSz = tx_of(W, H); e.txType = t; e.plane = p;
    e.dcQ = dc_q(I.qindex, I.bd); e.acQ = ac_q(I.qindex, I.bd);
    e.skipCtx = S.skipCtx[p]; e.dcCtx = S.dcCtx[p]; e.intraDir = p == 0 ? cand_mode(c) : 0; e.bd = I.bd;
    e.rnd = p ? P.rndC : P.rnd; e.zbin = p ? P.zbinC : P.zbin;
    e.resEnergy = 0.f;
    e.C = S.tab.cost;
    return half_eval<W, H>(res, tmp, lev, e, S.tab);
}

// one RD task (candidate c, 255 = CfL, of plane p; DCT_DCT) of a tx block with a side of 32 or 64 by the calling group
template <int W, int H>
__device__ EvalOut grp_task_big(lds<RcShared> &S, lds<RcTeam> &X, const Img &I, const DevTables &T, int p, int c,
                                const lds<uint16_t> *src, int sp, const SearchParams &P) {
    constexpr int LW = W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    const int lane = glane(), bd = I.bd;
    const int m = c == 255 ? DC_PRED : cand_mode(c), pa = pangle(m, c == 255 ? 0 : cand_delta(c));
    const lds<uint16_t> *EA = S.ea[p] + EDGE_OFF, *EL = S.el[p] + EDGE_OFF;
    int upA = 0, upL = 0;
    if (c != 255 && is_dir(m) && I.edgeFilter) {
        for (int i = lane; i < EDGE_LEN; i += GW) { X.ea[i] = S.ea[p][i]; X.el[i] = S.el[p][i]; }
        gsync();
        grp_edge_prepare(pa, LW, LH, S.ft[p], S.hvL[p], S.hvA[p], S.px[p], S.py[p], S.mx[p], S.my[p], X.ea + EDGE_OFF, X.el + EDGE_OFF, bd,
                         upA, upL);
        EA = X.ea + EDGE_OFF;
        EL = X.el + EDGE_OFF;
    }
    if (lane == 0) { S.gup[gid()][0] = upA; S.gup[gid()][1] = upL; }
    const int dcv = S.dcv[p], alpha = c == 255 ? S.cflAlpha[p - 1] : 0, cavg = S.cflAvg;
    auto res = [&](int i0, int j, auto &r) {
        constexpr int CH = sizeof(r) / sizeof(r[0]);
        if (c == 255) {
#pragma unroll
            for (int k = 0; k < CH; k++) r[k] = clip3(0, (1 << bd) - 1, dcv + round2signed(alpha * (S.cflL[(i0 + k) * W + j] - cavg), 6));
        } else {
            pred_col(m, pa, dcv, i0, 1, j, LW, LH, EA, EL, upA, upL, r);
        }
#pragma unroll
        for (int k = 0; k < CH; k++) r[k] = (int)src[(i0 + k) * sp + j] - r[k];
    };
    const int txSz = tx_of(W, H);
    RcEvalIn e;
    e.resEnergy = grp_fwd_dct<W, H>(X.tmp, T, S.tab.txScale[txSz], res);
    e.txSz = txSz; e.txType = DCT_DCT; e.plane = p;
    e.dcQ = dc_q(I.qindex, bd); e.acQ = ac_q(I.qindex, bd);
    e.skipCtx = S.skipCtx[p]; e.dcCtx = S.dcCtx[p]; e.intraDir = p == 0 ? cand_mode(c) : 0; e.bd = bd;
    e.rnd = p ? P.rndC : P.rnd; e.zbin = p ? P.zbinC : P.zbin;
    e.C = S.tab.cost;
    return grp_qrate<W, H>(X.tmp, W + 1, X.q, e, S.tab);
}

// ---- exact inverse transform + reconstruction by the workgroup (spec 7.13.3, 7.12.3 step 3) ----
template <int LW>
__device__ __forceinline__ uint32_t inv_row(int i, int W, int th, int tw, const lds<int16_t> *q, int dcQ, int acQ, int dsh, int bd, bool rect, int rk,
                            int rowShift, lds<int16_t> *res) {
    constexpr int N = 1 << LW;
    int32_t Tv[N];
#pragma unroll
    for (int j = 0; j < N; j++) {
        const int v = (i < th && j < tw) ? q[i * tw + j] : 0;
        Tv[j] = v ? dequant(v, (i == 0 && j == 0) ? dcQ : acQ, dsh, bd) : 0;
        if (rect) Tv[j] = Inv::rnd12(Inv::mul(Tv[j], 2896));   // |Tv| < 2^17: exact
    }
    uint32_t err = 0;
    Inv::run<LW>(Tv, rk, bd + 8, err);
    const int colR = bd + 6 > 16 ? bd + 6 : 16;
#pragma unroll
    for (int j = 0; j < N; j++) res[i * W + j] = (int16_t)clip3(-(1 << (colR - 1)), (1 << (colR - 1)) - 1, (int32_t)round2l(Tv[j], rowShift));
    return err;
}
template <int LH>
__device__ __forceinline__ uint32_t inv_col(int j, int W, int ck, int bd, lds<int16_t> *res) {
    constexpr int N = 1 << LH;
    int32_t Tv[N];
#pragma unroll
    for (int i = 0; i < N; i++) Tv[i] = res[i * W + j];
    uint32_t err = 0;
    Inv::run<LH>(Tv, ck, bd + 6 > 16 ? bd + 6 : 16, err);
#pragma unroll
    for (int i = 0; i < N; i++) res[i * W + j] = (int16_t)round2l(Tv[i], 4);
    return err;
}

// dequantize + inverse + add the levels q (LDS, modified by the repair) onto dst; returns the number of repairs
template <typename Px>
__device__ __forceinline__ int wg_inverse_add(lds<RcShared> &S, lds<int16_t> *q, int txSz, int txType, int bd, int dcQ, int acQ, Px *dst, int pitch) {
    const int tid = threadIdx.x, lw = av1t::Tx_Width_Log2[txSz], lh = av1t::Tx_Height_Log2[txSz], W = 1 << lw, H = 1 << lh;
    const int tw = W > 32 ? 32 : W, th = H > 32 ? 32 : H, n = tw * th, den = dq_denom(txSz), dsh = den == 4 ? 2 : den == 2 ? 1 : 0;
    const bool rect = lw - lh == 1 || lh - lw == 1;
    const int rk = row_kind(txType), ck = col_kind(txType), rs = av1t::Transform_Row_Shift[txSz];
    int fixes = 0;
    for (;;) {
        if (tid == 0) { S.errFlag = 0; S.nzFlag = 0; }
        __syncthreads();
        uint32_t err = 0;
        if (tid < H) {
            switch (lw) {
            case 2: err = inv_row<2>(tid, W, th, tw, q, dcQ, acQ, dsh, bd, rect, rk, rs, S.inv); break;
            case 3: err = inv_row<3>(tid, W, th, tw, q, dcQ, acQ, dsh, bd, rect, rk, rs, S.inv); break;
            case 4: err = inv_row<4>(tid, W, th, tw, q, dcQ, acQ, dsh, bd, rect, rk, rs, S.inv); break;
            case 5: err = inv_row<5>(tid, W, th, tw, q, dcQ, acQ, dsh, bd, rect, rk, rs, S.inv); break;
            default: err = inv_row<6>(tid, W, th, tw, q, dcQ, acQ, dsh, bd, rect, rk, rs, S.inv); break;
            }
        }
        if (err) atomicOr(gen(&S.errFlag), (int)err);
        __syncthreads();
        err = 0;
        if (tid < W) {
            switch (lh) {
            case 2: err = inv_col<2>(tid, W, ck, bd, S.inv); break;
            case 3: err = inv_col<3>(tid, W, ck, bd, S.inv); break;
            case 4: err = inv_col<4>(tid, W, ck, bd, S.inv); break;
            case 5: err = inv_col<5>(tid, W, ck, bd, S.inv); break;
            default: err = inv_col<6>(tid, W, ck, bd, S.inv); break;
            }
        }
        if (err) atomicOr(gen(&S.errFlag), (int)err);
        __syncthreads();
        if (!S.errFlag) break;
        fixes++;
        for (int i = tid; i < n; i += RC_NT) {
            q[i] = (int16_t)(q[i] / 2);
            if (q[i]) S.nzFlag = 1;
        }
        __syncthreads();
        if (!S.nzFlag) return fixes;
    }
    const int fud = flip_ud(txType), flr = flip_lr(txType), maxv = (1 << bd) - 1;
    for (int p = tid; p < W * H; p += RC_NT) {
        const int i = p >> lw, j = p & (W - 1);
        Px *d = dst + (size_t)(fud ? H - 1 - i : i) * pitch + (flr ? W - 1 - j : j);
        *d = (Px)clip3(0, maxv, (int)*d + S.inv[i * W + j]);
    }
    __syncthreads();
    return fixes;
}

// write the prediction of candidate c of plane p (edges prepared by group g, or the base edges) into dst (workgroup)
template <typename Px>
__device__ void wg_write_pred(const lds<RcShared> &S, const lds<uint16_t> *gea, const lds<uint16_t> *gel, int g, int p, int c, int lw, int lh, Px *dst,
                              int pitch, int bd) {
    const int W = 1 << lw, H = 1 << lh;
    const int m = c == 255 ? DC_PRED : cand_mode(c), pa = pangle(m, c == 255 ? 0 : cand_delta(c));
    const bool filt = c != 255 && is_dir(m) && gea != nullptr;
    const lds<uint16_t> *EA = filt ? gea + EDGE_OFF : S.ea[p] + EDGE_OFF, *EL = filt ? gel + EDGE_OFF : S.el[p] + EDGE_OFF;
    const int upA = filt ? S.gup[g][0] : 0, upL = filt ? S.gup[g][1] : 0, dcv = S.dcv[p];
    for (int q = threadIdx.x; q < W * H; q += RC_NT) {
        const int i = q >> lw, j = q & (W - 1);
        int pr;
        if (c == 255) pr = clip3(0, (1 << bd) - 1, dcv + round2signed(S.cflAlpha[p - 1] * (S.cflL[i * W + j] - S.cflAvg), 6));
        else if (m == DC_PRED) pr = dcv;
        else pr = pred_px(m, pa, i, j, lw, lh, EA, EL, upA, upL);
        dst[(size_t)i * pitch + j] = (Px)pr;
    }
    __syncthreads();
}

// record the finished tx block with levels q: coefficients, txb entry, context maps, BlockDecoded (workgroup)
__device__ int wg_record_txb(lds<RcShared> &S, const Batch &B, uint32_t sb, const Img &I, int p, int txSz, int txType, int x4, int y4,
                             const lds<int16_t> *q) {
    const int tid = threadIdx.x, n = tx_ncoef(txSz);
    const uint32_t off = (uint32_t)S.nCoef;
    if (tid < 32) {
        // eob: 1 + the largest scan index of a nonzero level (inverse default scan; rows for class VERT, columns for
        // HORIZ); cul: the sum of |level| before the eob, i.e. over every level (later ones are 0), capped at 63
        const int cls = tx_class(txType), tw = tx_w(txSz) > 32 ? 32 : tx_w(txSz), th = tx_h(txSz) > 32 ? 32 : tx_h(txSz);
        const uint16_t *isc = S.tab.iscan + dscan_off(av1t::Adjusted_Tx_Size[txSz]);
        int last = 0, cul = 0;
        for (int i = tid; i < n; i += 32) {
            const int v = q[i];
            if (v) {
                const int c = cls == TX_CLASS_2D ? (int)isc[i] : cls == TX_CLASS_VERT ? i : (i % tw) * th + i / tw;
                last = max(last, c + 1);
                cul += abs(v);
            }
        }
        const int eob = gmaxi(last);
        cul = gsumi(cul);
        if (tid == 0) { S.eobOut = eob; S.culOut = min(cul, 63); }
    }
    int16_t *oc = B.oCoeffs + (size_t)sb * B.coefCap + off;
    for (int i = tid; i < n; i += RC_NT) oc[i] = q[i];
    __syncthreads();
    const int eob = S.eobOut, cul = S.culOut, dcCat = q[0] < 0 ? 1 : q[0] > 0 ? 2 : 0;
    const int ss = p ? I.ss : 0, stride = I.MiCols >> ss, rows = I.MiRows >> ss;
    const int w4 = tx_w(txSz) >> 2, h4 = tx_h(txSz) >> 2, lw4 = av1t::Tx_Width_Log2[txSz] - 2;
    const int r4 = y4 & ((16 >> ss) - 1), c4 = x4 & ((16 >> ss) - 1);
    for (int k = tid; k < w4 * h4; k += RC_NT) {
        const int yy = k >> lw4, xx = k & (w4 - 1);
        if (y4 + yy < rows && x4 + xx < stride) {
            I.cLvl[p][(size_t)(y4 + yy) * stride + x4 + xx] = (uint8_t)(eob ? cul : 0);
            I.cDc[p][(size_t)(y4 + yy) * stride + x4 + xx] = (uint8_t)(eob ? dcCat : 0);
        }
        S.bdm[p][r4 + yy + 1][c4 + xx + 1] = 1;
    }
    if (tid == 0) {
        RecTxb t;
        t.coeff_off = off; t.x4 = (uint16_t)x4; t.y4 = (uint16_t)y4; t.plane = (uint8_t)p; t.tx_size = (uint8_t)txSz;
        t.tx_type = (uint8_t)(p == 0 && eob == 0 ? DCT_DCT : txType); t.flags = 0; t.eob = (uint16_t)eob; t.reserved = 0;
        B.oTxb[(size_t)sb * B.txbCap + S.nTxb] = t;
    }
    __syncthreads();
    if (tid == 0) { S.nTxb++; S.nCoef += n; }
    __syncthreads();
    return eob;
}

// source of plane p in the staging buffer
__device__ __forceinline__ const lds<uint16_t> *stage_src(const lds<RcShared> &S, int p) { return S.srcb + (p == 2 ? 2048 : 0); }

// Inlined into each per-shape leaf function (rc_luma / rc_chroma): as a call of its own it would run on the ~140
// registers outside the callee-saved ones and spill; inlined, the leaf function's one frame covers it.
// run S.taskC/T/P[0 .. nTask) on the 8 groups (the source block staged in S.srcb, row stride W); per task
// J = fmaf(lambda, R + extra, D) with extra = luma mode bits (plane 0) or 0; S.taskJ / taskE.  Then S.win = the first
// minimum over the tasks.  Blocks up to 16x16: luma tasks read the residual of their candidate (S.taskK), made by
// wg_luma_residuals; chroma tasks make their own in slot t; levels of task t go to S.u.s.lev[t].  Larger blocks:
// team g (tasks g, g + nTeam, ..., run in descending order) keeps the levels and prepared edges of task g.
template <int W, int H>
__device__ __forceinline__ void wg_run_tasks(lds<RcShared> &S, const Img &I, const DevTables &T, const SearchParams &P, float lambda, bool lumaBits,
                             bool pickWinner) {
    constexpr bool big = W > 16 || H > 16;
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 4 ? 2 : H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    if constexpr (!big) {   // the packed levels hold every task: 6 luma candidates x the set's types, or 14 chroma tasks
        constexpr int set = tx_set_intra(tx_of(W, H));
        constexpr int nt = set == TX_SET_INTRA_1 ? 7 : set == TX_SET_INTRA_2 ? 5 : 1;
        static_assert((6 * nt > 14 ? 6 * nt : 14) * W * H <= RC_LEV_TOTAL, "the packed levels cannot hold every task");
    }
    const int g = gid(), lane = glane();
    const CostTab &C = *S.tab.cost;
    // blocks of up to 64 coefficients: 16 half groups, two tasks per group at once; luma residuals are ready, chroma
    // tasks first make theirs (one task per group, as below), then the halves evaluate
    bool done = false;
    if constexpr (!big && W * H <= 64) {
        done = true;
        if (!lumaBits) {
            for (int t = g; t < S.nTask; t += RC_NG)
                grp_residual<W, H>(S, I.edgeFilter ? S.u.s.g[g].ea : nullptr, S.u.s.g[g].el, S.taskP[t], S.taskC[t], stage_src(S, S.taskP[t]), W,
                             S.u.s.res[t], I.bd);
            __syncthreads();
        }
        const int h = lane >> 4;
        for (int t = 2 * g + h; t < S.nTask; t += 2 * RC_NG) {
            const int c = S.taskC[t], tt = S.taskT[t], p = S.taskP[t];
            const EvalOut o = half_task<W, H>(S, S.u.s.g[g].tmp + h * 136, I, p, c, tt, S.u.s.res[lumaBits ? S.taskK[t] : t], S.u.s.lev + t * (W * H), P);
            float J;
            if (lumaBits) {
                const int m = cand_mode(c);
                float mb = C.y_mode[S.am][S.lm][m];
                if (S.bsize >= BLOCK_8X8 && is_dir(m)) mb += C.angle_delta[m - V_PRED][cand_delta(c) + 3];
                J = fmaf(lambda, o.R + mb, o.D);
            } else {
                J = fmaf(lambda, o.R, o.D);
            }
            if ((lane & 15) == 0) { S.taskJ[t] = J; S.taskE[t] = o.eob; }
            gsync();
        }
    }
    const int NGT = big ? S.nTeam : RC_NG;   // big tasks: the teams (groups 0 .. nTeam - 1)
    // a team keeps the levels and prepared edges of its last task only: it runs its tasks in descending order, so the
    // first tasks -- the analysis's best candidates, the likeliest winners -- are the ones kept
    const int tFirst = big ? (g < S.nTask ? g + ((S.nTask - 1 - g) / NGT) * NGT : -1) : g, tStep = big ? -NGT : NGT;
    if (!done && g < NGT) for (int t = tFirst; t >= 0 && t < S.nTask; t += tStep) {
        const int c = S.taskC[t], tt = S.taskT[t], p = S.taskP[t];
        EvalOut o;
        if constexpr (big) {
            o = grp_task_big<W, H>(S, team_scr(S, g), I, T, p, c, stage_src(S, p), W, P);
        } else {
            if (!lumaBits)   // chroma: this task's own residual
                grp_residual<W, H>(S, I.edgeFilter ? S.u.s.g[g].ea : nullptr, S.u.s.g[g].el, p, c, stage_src(S, p), W, S.u.s.res[t], I.bd);
            o = grp_task<W, H>(S, S.u.s.g[g], I, p, c, tt, S.u.s.res[lumaBits ? S.taskK[t] : t], S.u.s.lev + t * (W * H), P);
        }
        float J;
        if (lumaBits) {
            const int m = cand_mode(c);
            float mb = C.y_mode[S.am][S.lm][m];
            if (S.bsize >= BLOCK_8X8 && is_dir(m)) mb += C.angle_delta[m - V_PRED][cand_delta(c) + 3];
            J = fmaf(lambda, o.R + mb, o.D);
        } else {
            J = fmaf(lambda, o.R, o.D);
        }
        if (lane == 0) { S.taskJ[t] = J; S.taskE[t] = o.eob; }
        gsync();
    }
    __syncthreads();
    if (S.flopsOn && threadIdx.x < GW) {   // work accounting: the forward transforms of the tasks
        int f = 0;
        for (int t = lane; t < S.nTask; t += GW) f += (int)fwd_flops(W, H, S.taskT[t]);
        f = gsumi(f);
        if (lane == 0) S.flops += (unsigned long long)f;
    }
    if (pickWinner && threadIdx.x < GW) {   // the first minimum over the tasks (at most 64)
        const int n = S.nTask;
        float v = lane < n ? S.taskJ[lane] : INF;
        int i = lane;
        if (lane + GW < n && S.taskJ[lane + GW] < v) { v = S.taskJ[lane + GW]; i = lane + GW; }
        const int w = gargmin(v, i);
        if (lane == 0) S.win = w < 0 ? 0 : w;
    }
    __syncthreads();
}

// luma blocks up to 16x16: the residual of each candidate S.candC[k] into S.u.s.res[k] (group k)
template <int W, int H>
__device__ void wg_luma_residuals(lds<RcShared> &S, const Img &I) {
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : 4, LH = H == 4 ? 2 : H == 8 ? 3 : 4;
    const int g = gid();
    if (g < S.nCand)
        grp_residual<W, H>(S, I.edgeFilter ? S.u.s.g[g].ea : nullptr, S.u.s.g[g].el, 0, S.candC[g], stage_src(S, 0), W, S.u.s.res[g], I.bd);
    __syncthreads();
}

// the levels of the winning task t, and (larger blocks) the team holding them with its prepared edges: the team's
// kept task is t (t < nTeam), or it evaluates t again
template <int W, int H>
__device__ __forceinline__ lds<int16_t> *wg_final_levels(lds<RcShared> &S, const Img &I, const DevTables &T, const SearchParams &P, int t, int &team) {
    team = (W > 16 || H > 16) ? t % S.nTeam : t % RC_NG;
    if constexpr (W > 16 || H > 16) {
        if (t >= S.nTeam) {   // overwritten by a task of the same team that ran after it
            const int p = S.taskP[t];
            if (gid() == team) grp_task_big<W, H>(S, team_scr(S, team), I, T, p, S.taskC[t], stage_src(S, p), W, P);
            __syncthreads();
        }
        return team_scr(S, team).q;
    } else {
        return S.u.s.lev + t * (W * H);
    }
}

// write the prediction of task t of plane p (into dst): larger blocks predict again from team g's prepared edges;
// blocks up to 16x16 take source minus the kept residual (slot k)
template <int W, int H, typename Px>
__device__ void wg_final_pred(const lds<RcShared> &S, const Img &I, int p, int t, int k, int team, Px *dst, int pitch) {
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 4 ? 2 : H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    if constexpr (W > 16 || H > 16) {
        wg_write_pred<Px>(S, I.edgeFilter ? team_scr(S, team).ea : nullptr, team_scr(S, team).el, team, p, S.taskC[t], LW, LH, dst, pitch, I.bd);
    } else {
        const lds<uint16_t> *src = stage_src(S, p);
        const lds<int16_t> *res = S.u.s.res[k];
        for (int q = threadIdx.x; q < W * H; q += RC_NT) dst[(size_t)(q >> LW) * pitch + (q & (W - 1))] = (Px)((int)src[q] - res[q]);
        __syncthreads();
    }
}

template <int W, int H, typename Px>
__device__ void rc_luma(lds<RcShared> &S, const Batch &B, const SearchParams &P, uint32_t sb, const Img &I, int bsize, int x, int y, int availU,
                        int availL, int tx0, int ty0, float lambda) {
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 4 ? 2 : H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    const DevTables &T = *B.T;
    const int tid = threadIdx.x, txSz = tx_of(W, H);
    const int sbr4 = (y >> 2) & 15, sbc4 = (x >> 2) & 15, mi_row = y >> 2, mi_col = x >> 2;
    {
        const Px *src = src_plane<Px>(I, 0) + (size_t)y * I.spitch[0] + x;
        for (int q = tid; q < W * H; q += RC_NT) S.srcb[q] = src[(size_t)(q >> LW) * I.spitch[0] + (q & (W - 1))];
    }
    if (gid() == 0)
        rc_edges<Px>(S, I, 0, x, y, LW, LH, availL, availU, S.bdm[0][sbr4][sbc4 + (W >> 2) + 1], S.bdm[0][sbr4 + (H >> 2) + 1][sbc4],
                     mi_row, mi_col);
    if (gid() == 2) grp_txb_ctx(I, 0, txSz, mi_col, mi_row, bsize, tx0, ty0, &S.skipCtx[0], &S.dcCtx[0]);
    if (tid >= 128 && tid < 192) {   // wave 2: the tasks (candidate k, tx type t), k-major, types ascending
        const int w = tid - 128;
        const BlockRes &br = B.blk[I.shOff[bsize] + (uint32_t)(y / (4 * bh4(bsize))) * I.shNbx[bsize] + x / (4 * bw4(bsize))];
        const int nc = max(1, min((int)br.ncand, P.Krecheck));
        const int set = tx_set_intra(txSz);
        if (w < nc) S.candC[w] = br.cand[w];
        if (w == 0) {
            S.am = av1t::Intra_Mode_Context[availU ? I.mY[(size_t)(mi_row - 1) * I.MiCols + mi_col] : DC_PRED];
            S.lm = av1t::Intra_Mode_Context[availL ? I.mY[(size_t)mi_row * I.MiCols + mi_col - 1] : DC_PRED];
            S.nCand = nc;
            S.bsize = bsize;
        }
        if (P.txAllCL && set != TX_SET_DCTONLY) {
            // every pair at once: with nt types in the set, pair (k, t) is task k * nt + (types of the set below t)
            uint32_t sm = 0;
            for (int t = 0; t < 16; t++) sm |= tx_in_set(set, t) ? 1u << t : 0u;
            const int nt = __popc(sm);
            for (int i = w; i < nc * 16; i += 64) {
                const int k = i >> 4, t = i & 15;
                if ((sm >> t) & 1) {
                    const int n = k * nt + __popc(sm & ((1u << t) - 1));
                    S.taskC[n] = br.cand[k]; S.taskT[n] = t; S.taskP[n] = 0; S.taskK[n] = k;
                }
            }
            if (w == 0) S.nTask = nc * nt;
        } else if (w == 0) {
            int n = 0;
            for (int k = 0; k < nc; k++) { S.taskC[n] = br.cand[k]; S.taskT[n] = DCT_DCT; S.taskP[n] = 0; S.taskK[n] = k; n++; }
            if (br.tx != DCT_DCT) { S.taskC[n] = br.cand[0]; S.taskT[n] = br.tx; S.taskP[n] = 0; S.taskK[n] = 0; n++; }
            S.nTask = n;
        }
    }
    __syncthreads();
    if constexpr (W <= 16 && H <= 16) wg_luma_residuals<W, H>(S, I);
    wg_run_tasks<W, H>(S, I, T, P, lambda, true, true);
    if (!(P.txAllCL && tx_set_intra(txSz) != TX_SET_DCTONLY) && P.txSearchCL && tx_set_intra(txSz) != TX_SET_DCTONLY) {
        // tx-type pass for the pass-1 winner (txAllCL off): the other types of the set
        if (tid == 0) {
            const int c0 = S.taskC[S.win], k0 = S.taskK[S.win], n0 = S.nTask;
            const BlockRes &br = B.blk[I.shOff[bsize] + (uint32_t)(y / (4 * bh4(bsize))) * I.shNbx[bsize] + x / (4 * bw4(bsize))];
            int n = n0;
            for (int t = 1; t < 16; t++)
                if (tx_in_set(tx_set_intra(txSz), t) && !(c0 == br.cand[0] && t == br.tx)) {
                    S.taskC[n] = c0; S.taskT[n] = t; S.taskP[n] = 0; S.taskK[n] = k0; n++;
                }
            S.nTask = n;
        }
        __syncthreads();
        wg_run_tasks<W, H>(S, I, T, P, lambda, true, true);
    }
    const int win = S.win;
    int team;
    lds<int16_t> *q = wg_final_levels<W, H>(S, I, T, P, win, team);
    Px *dst = rec_plane<Px>(I, 0) + (size_t)y * I.pitch[0] + x;
    wg_final_pred<W, H, Px>(S, I, 0, win, S.taskK[win], team, dst, I.pitch[0]);
    if (S.taskE[win] > 0) {
        const int fixes = wg_inverse_add<Px>(S, q, txSz, S.taskT[win], I.bd, dc_q(I.qindex, I.bd), ac_q(I.qindex, I.bd), dst, I.pitch[0]);
        if (fixes && tid == 0) atomicAdd(&B.stats[3], 1u);
    }
    const int eob = wg_record_txb(S, B, sb, I, 0, txSz, S.taskT[win], x >> 2, y >> 2, q);
    if (tid == 0) { S.mode = cand_mode(S.taskC[win]); S.delta = cand_delta(S.taskC[win]); S.winEob[0] = eob; }
    __syncthreads();
}

template <int W, int H, typename Px>
__device__ void rc_chroma(lds<RcShared> &S, const Batch &B, const SearchParams &P, uint32_t sb, const Img &I, int bsize, int mi_row, int mi_col,
                          int availUC, int availLC, int tx0, int ty0, float lambda) {
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 4 ? 2 : H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    const DevTables &T = *B.T;
    const CostTab &C = *S.tab.cost;
    const int tid = threadIdx.x, ss = I.ss, bd = I.bd, txSz = tx_of(W, H);
    const int cx = (mi_col >> ss) * 4, cy = (mi_row >> ss) * 4, x4 = cx >> 2, y4 = cy >> 2;
    const int sbr4 = (mi_row & 15) >> ss, sbc4 = (mi_col & 15) >> ss;
    if (gid() < 2) {
        const int p = 1 + gid();
        rc_edges<Px>(S, I, p, cx, cy, LW, LH, availLC, availUC, S.bdm[p][sbr4][sbc4 + (W >> 2) + 1], S.bdm[p][sbr4 + (H >> 2) + 1][sbc4],
                     mi_row, mi_col);
    }
    if (gid() == 2) grp_txb_ctx(I, 1, txSz, x4, y4, bsize, tx0, ty0, &S.skipCtx[1], &S.dcCtx[1]);
    if (gid() == 3) grp_txb_ctx(I, 2, txSz, x4, y4, bsize, tx0, ty0, &S.skipCtx[2], &S.dcCtx[2]);
    {
        const Px *gU = src_plane<Px>(I, 1) + (size_t)cy * I.spitch[1] + cx, *gV = src_plane<Px>(I, 2) + (size_t)cy * I.spitch[2] + cx;
        for (int q = tid; q < W * H; q += RC_NT) {
            S.srcb[q] = gU[(size_t)(q >> LW) * I.spitch[1] + (q & (W - 1))];
            S.srcb[2048 + q] = gV[(size_t)(q >> LW) * I.spitch[2] + (q & (W - 1))];
        }
    }
    if (tid < 32) S.satdI[tid] = 0;
    if (tid < 8) S.acc64[tid] = 0;
    __syncthreads();
    constexpr int sp = W;
    const lds<uint16_t> *srcU = S.srcb, *srcV = S.srcb + 2048;
    const bool cflOk = P.cfl && cfl_allowed(bsize);
    const float lamS = P.satdLambda * T.lamS[bd == 10][I.qindex];   // the luma lambda's screening weight
    const int yMode = S.mode;
    auto uvbits = [&](int m) -> float { return cfl_allowed(bsize) ? C.uv_cfl_yes[yMode][m] : C.uv_cfl_no[yMode][m < 13 ? m : 0]; };
    auto uvcbits = [&](int c) -> float {
        const int m = cand_mode(c);
        float b = uvbits(m);
        if (bsize >= BLOCK_8X8 && is_dir(m)) b += C.angle_delta[m - V_PRED][cand_delta(c) + 3];
        return b;
    };
    // ---- CfL: luma values, average, least-squares alpha, +-1 refinement by exact SSE ----
    if constexpr (W * H <= 64) if (cflOk) {
        // up to 64 samples: wave 0 alone, one sample per lane, with 64-lane integer sums (the same integers as the
        // workgroup form below) and no barrier until the result
        if (tid < 64) {
            const int lt = av1t::Max_Tx_Size_Rect[bsize], tw = tx_w(lt), th = tx_h(lt);
            int maxLumaW = 0, maxLumaH = 0;
            for (int ly = mi_row * 4; ly < mi_row * 4 + 4 * bh4(bsize); ly += th)
                for (int lx = mi_col * 4; lx < mi_col * 4 + 4 * bw4(bsize); lx += tw)
                    if (lx < I.MiCols * 4 && ly < I.MiRows * 4) { maxLumaW = lx + tw; maxLumaH = ly + th; }
            const Px *luma = rec_plane<Px>(I, 0);
            const int lp = I.pitch[0], i = tid >> LW, j = tid & (W - 1);
            const bool in = tid < W * H;
            int v = 0;
            if (in) {
                if (ss) {
                    const int ly = min((cy + i) << 1, maxLumaH - 2), lx = min((cx + j) << 1, maxLumaW - 2);
                    v = ((int)luma[(size_t)ly * lp + lx] + luma[(size_t)ly * lp + lx + 1] + luma[(size_t)(ly + 1) * lp + lx] + luma[(size_t)(ly + 1) * lp + lx + 1]) << 1;
                } else {
                    v = (int)luma[(size_t)min(cy + i, maxLumaH - 1) * lp + min(cx + j, maxLumaW - 1)] << 3;
                }
                S.cflL[tid] = v;
            }
            int sum = v;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
            const int avg = round2(sum, LW + LH);
            const long long l = in ? (long long)(v - avg) : 0;
            long long nU = in ? l * ((int)srcU[i * sp + j] - S.dcv[1]) : 0, nV = in ? l * ((int)srcV[i * sp + j] - S.dcv[2]) : 0, dd = l * l;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                nU += __shfl_xor(nU, o, 64);
                nV += __shfl_xor(nV, o, 64);
                dd += __shfl_xor(dd, o, 64);
            }
            int alpha[2];
            for (int k = 0; k < 2; k++) {
                const long long n = k == 0 ? nU : nV;
                const int a = dd > 0 ? (int)rintf(64.f * (float)n / (float)dd) : 0;
                alpha[k] = max(-16, min(16, a));
            }
            long long sse[6] = {0, 0, 0, 0, 0, 0};
            if (in) {
#pragma unroll
                for (int pl = 0; pl < 2; pl++) {
                    const int s = pl == 0 ? srcU[i * sp + j] : srcV[i * sp + j];
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const int a = max(-16, min(16, alpha[pl] + k - 1));
                        const int d = s - clip3(0, (1 << bd) - 1, S.dcv[pl + 1] + round2signed(a * (int)l, 6));
                        sse[pl * 3 + k] = (long long)d * d;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 6; k++)
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) sse[k] += __shfl_xor(sse[k], o, 64);
            if (tid == 0) {
                S.cflAvg = avg;
                for (int pl = 0; pl < 2; pl++) {
                    int bk = 1;
                    for (int k = 0; k < 3; k++) if (sse[pl * 3 + k] < sse[pl * 3 + bk]) bk = k;
                    S.cflAlpha[pl] = max(-16, min(16, alpha[pl] + bk - 1));
                }
            }
        }
        if (tid < 32) S.satdI[tid] = 0;
        __syncthreads();
    } else if (tid == 0) {
        S.cflAlpha[0] = S.cflAlpha[1] = 0;
    }
    if constexpr (W * H > 64) if (cflOk) {
        const int lt = av1t::Max_Tx_Size_Rect[bsize], tw = tx_w(lt), th = tx_h(lt);
        int maxLumaW = 0, maxLumaH = 0;
        for (int ly = mi_row * 4; ly < mi_row * 4 + 4 * bh4(bsize); ly += th)
            for (int lx = mi_col * 4; lx < mi_col * 4 + 4 * bw4(bsize); lx += tw)
                if (lx < I.MiCols * 4 && ly < I.MiRows * 4) { maxLumaW = lx + tw; maxLumaH = ly + th; }
        const Px *luma = rec_plane<Px>(I, 0);
        const int lp = I.pitch[0];
        int sum = 0;
        for (int q = tid; q < W * H; q += RC_NT) {
            const int i = q >> LW, j = q & (W - 1);
            int v;
            if (ss) {
                const int ly = min((cy + i) << 1, maxLumaH - 2), lx = min((cx + j) << 1, maxLumaW - 2);
                v = ((int)luma[(size_t)ly * lp + lx] + luma[(size_t)ly * lp + lx + 1] + luma[(size_t)(ly + 1) * lp + lx] + luma[(size_t)(ly + 1) * lp + lx + 1]) << 1;
            } else {
                v = (int)luma[(size_t)min(cy + i, maxLumaH - 1) * lp + min(cx + j, maxLumaW - 1)] << 3;
            }
            S.cflL[q] = v;
            sum += v;
        }
        sum = gsumi(sum);
        if (glane() == 0) atomicAdd(gen(&S.satdI[31]), sum);
        __syncthreads();
        if (tid == 0) S.cflAvg = round2(S.satdI[31], LW + LH);
        __syncthreads();
        long long nU = 0, nV = 0, dd = 0;
        for (int q = tid; q < W * H; q += RC_NT) {
            const int i = q >> LW, j = q & (W - 1);
            const long long l = S.cflL[q] - S.cflAvg;
            nU += l * ((int)srcU[i * sp + j] - S.dcv[1]);
            nV += l * ((int)srcV[i * sp + j] - S.dcv[2]);
            dd += l * l;
        }
        nU = gsuml(nU); nV = gsuml(nV); dd = gsuml(dd);
        if (glane() == 0) {
            atomicAdd(gen(&S.acc64[0]), (unsigned long long)nU);
            atomicAdd(gen(&S.acc64[1]), (unsigned long long)nV);
            atomicAdd(gen(&S.acc64[2]), (unsigned long long)dd);
        }
        __syncthreads();
        if (tid == 0) {
            const long long d = (long long)S.acc64[2];
            for (int k = 0; k < 2; k++) {
                const long long n = (long long)S.acc64[k];
                const int a = d > 0 ? (int)rintf(64.f * (float)n / (float)d) : 0;
                S.cflAlpha[k] = max(-16, min(16, a));
            }
            for (int k = 0; k < 6; k++) S.acc64[2 + k] = 0;
        }
        __syncthreads();
        long long sse[6] = {0, 0, 0, 0, 0, 0};
        for (int q = tid; q < W * H; q += RC_NT) {
            const int i = q >> LW, j = q & (W - 1), l = S.cflL[q] - S.cflAvg;
#pragma unroll
            for (int pl = 0; pl < 2; pl++) {
                const int s = pl == 0 ? srcU[i * sp + j] : srcV[i * sp + j];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int a = max(-16, min(16, S.cflAlpha[pl] + k - 1));
                    const int d = s - clip3(0, (1 << bd) - 1, S.dcv[pl + 1] + round2signed(a * l, 6));
                    sse[pl * 3 + k] += (long long)d * d;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const long long v = gsuml(sse[k]);
            if (glane() == 0) atomicAdd(gen(&S.acc64[2 + k]), (unsigned long long)v);
        }
        __syncthreads();
        if (tid == 0)
            for (int pl = 0; pl < 2; pl++) {
                int bk = 1;
                for (int k = 0; k < 3; k++) if ((long long)S.acc64[2 + pl * 3 + k] < (long long)S.acc64[2 + pl * 3 + bk]) bk = k;
                S.cflAlpha[pl] = max(-16, min(16, S.cflAlpha[pl] + bk - 1));
            }
        if (tid < 32) S.satdI[tid] = 0;
        __syncthreads();
    } else if (tid == 0) {
        S.cflAlpha[0] = S.cflAlpha[1] = 0;
    }
    // ---- SATD screening of the UV candidates over U and V (unfiltered edges) ----
    if (tid < 13) S.uvc[tid] = cand_pack(tid, 0);
    __syncthreads();
    constexpr int NSUB = (W / 4) * (H / 4), NSX = W / 4;
    // screening of blocks of up to 16 4x4s: item = (candidate, plane, pixel), 16 lanes per 4x4 and the Hadamard
    // across them, so a wave holds at most two candidates (modes); larger blocks: item = (candidate, plane, 4x4).  A
    // candidate's items are consecutive, so each segment of up to 32 lanes sums its costs with shuffles before one LDS
    // atomic.
    auto screen = [&](int c0, int c1) {
        constexpr int NSY = NSUB / NSX;
        if constexpr (NSUB <= 16 && NSY >= 2) {
            // item = (candidate, plane, 4x4 column, pixel of the 4x4): each thread predicts a vertical run of pred_col
            // through the column's NSY 4x4s at its (row, column) = ((t >> 2) & 3, t & 3); each 4x4's Hadamard runs
            // across the 16 lanes holding it; a (candidate, plane) spans NSX x 16 consecutive threads (a whole wave
            // for 16-wide blocks), so a round holds few modes
            constexpr int TPC = NSX * 16, SEG = TPC < GW ? TPC : GW;
            const int items = (c1 - c0) * 2 * TPC;
            for (int base = 0; base < items; base += RC_NT) {
                const int it = base + tid;
                int d[NSY], ci = c0;
#pragma unroll
                for (int k = 0; k < NSY; k++) d[k] = 0;
                if (it < items) {
                    ci = c0 + it / (2 * TPC);
                    const int rr = it % (2 * TPC), p = 1 + rr / TPC, q = rr % TPC;
                    const int sx = (q >> 4) * 4 + (q & 3), sy0 = (q >> 2) & 3;
                    const int c = S.uvc[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
                    const lds<uint16_t> *s = p == 1 ? srcU : srcV;
                    const lds<uint16_t> *A = S.ea[p] + EDGE_OFF, *L = S.el[p] + EDGE_OFF;
                    int pr[NSY];
                    pred_col(m, pa, S.dcv[p], sy0, 4, sx, LW, LH, A, L, 0, 0, pr);
#pragma unroll
                    for (int k = 0; k < NSY; k++) d[k] = (int)s[(sy0 + 4 * k) * sp + sx] - pr[k];
                }
                int v = 0;
#pragma unroll
                for (int k = 0; k < NSY; k++) v += hadamard4_lanes(d[k]);
                v = seg_sumi<SEG>(v);
                if (it < items && (it & (SEG - 1)) == 0) atomicAdd(gen(&S.satdI[ci]), v);
            }
            __syncthreads();
            return;
        }
        if constexpr (NSUB <= 16) {
            constexpr int PIX = NSUB * 16;
            const int items = (c1 - c0) * PIX * 2;
            for (int base = 0; base < items; base += RC_NT) {
                const int it = base + tid;
                int v = 0, ci = c0;
                if (it < items) {
                    ci = c0 + it / (2 * PIX);
                    const int rr = it % (2 * PIX), p = 1 + rr / PIX, px = rr % PIX, sbk = px >> 4;
                    const int sx = (sbk % NSX) * 4 + (px & 3), sy = (sbk / NSX) * 4 + ((px >> 2) & 3);
                    const int c = S.uvc[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
                    const lds<uint16_t> *s = p == 1 ? srcU : srcV;
                    const lds<uint16_t> *A = S.ea[p] + EDGE_OFF, *L = S.el[p] + EDGE_OFF;
                    const int pr = m == DC_PRED ? S.dcv[p] : pred_px(m, pa, sy, sx, LW, LH, A, L, 0, 0);
                    v = (int)s[sy * sp + sx] - pr;
                }
                v = seg_sumi<GW>(hadamard4_lanes(v));
                if (it < items && (it & (GW - 1)) == 0) atomicAdd(gen(&S.satdI[ci]), v);
            }
            __syncthreads();
            return;
        }
        constexpr int SEG = 2 * NSUB < GW ? 2 * NSUB : GW;
        const int items = (c1 - c0) * NSUB * 2;
        for (int base = 0; base < items; base += RC_NT) {
            const int it = base + tid;
            int v = 0, ci = c0;
            if (it < items) {
                ci = c0 + it / (2 * NSUB);
                const int rr = it % (2 * NSUB), p = 1 + rr / NSUB, sbk = rr % NSUB;
                const int sx = (sbk % NSX) * 4, sy = (sbk / NSX) * 4, c = S.uvc[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
                const lds<uint16_t> *s = p == 1 ? srcU : srcV;
                const lds<uint16_t> *A = S.ea[p] + EDGE_OFF, *L = S.el[p] + EDGE_OFF;
                int d[16];
#pragma unroll
                for (int j = 0; j < 4; j++) {   // one run of pred_col per column of the 4x4
                    int r[4];
                    pred_col(m, pa, S.dcv[p], sy, 1, sx + j, LW, LH, A, L, 0, 0, r);
#pragma unroll
                    for (int i = 0; i < 4; i++) d[i * 4 + j] = (int)s[(sy + i) * sp + sx + j] - r[i];
                }
                v = hadamard4(d);
            }
            v = seg_sumi<SEG>(v);
            if (it < items && (it & (SEG - 1)) == 0) atomicAdd(gen(&S.satdI[ci]), v);
        }
        __syncthreads();
    };
    screen(0, 13);
    if (tid < GW) {   // the deltas around the best directional mode (first minimum of the screening cost)
        const bool dl = P.uvDeltas && bsize >= BLOCK_8X8;
        if (dl) {
            const float j = tid < 8 ? fmaf(lamS, uvcbits(cand_pack(V_PRED + tid, 0)), (float)S.satdI[V_PRED + tid] * 0.25f) : INF;
            const int bm = V_PRED + gargmin(j, tid);
            if (tid < 6) S.uvc[13 + tid] = cand_pack(bm, tid < 3 ? -(tid + 1) : tid - 2);
        }
        if (tid == 0) S.nuv = dl ? 19 : 13;
    }
    __syncthreads();
    if (S.nuv > 13) screen(13, 19);
    if (tid < GW) {   // the uvK candidates of smallest screening cost (first minimum each round), U and V tasks each
        // taken at once: a candidate's place in that order is its rank under (cost, index), so the ones of rank < K
        // (finite cost) are the rounds' picks and rank k fills tasks 2k (U) and 2k + 1 (V)
        const int l = tid, nuv = S.nuv, K = max(1, min(P.uvK, 6));
        const float js = l < nuv ? fmaf(lamS, uvcbits(S.uvc[l]), (float)S.satdI[l] * 0.25f) : INF;
        S.uvJ[l] = js;
        gsync();
        int rank = 0, fin = 0;
        for (int j = 0; j < nuv; j++) {
            const float o = S.uvJ[j];
            rank += (o < js || (o == js && j < l)) ? 1 : 0;
            fin += o < INF ? 1 : 0;
        }
        if (js < INF && rank < K) {
            const int c = S.uvc[l], tt = chroma_tx_type(txSz, cand_mode(c));
            S.taskC[2 * rank] = c; S.taskT[2 * rank] = tt; S.taskP[2 * rank] = 1;
            S.taskC[2 * rank + 1] = c; S.taskT[2 * rank + 1] = tt; S.taskP[2 * rank + 1] = 2;
        }
        int n = 2 * min(K, fin);
        if (cflOk && (S.cflAlpha[0] || S.cflAlpha[1])) {
            if (l < 2) { S.taskC[n + l] = 255; S.taskT[n + l] = chroma_tx_type(txSz, UV_CFL_PRED); S.taskP[n + l] = 1 + l; }
            n += 2;
        }
        if (l == 0) S.nTask = n;
    }
    __syncthreads();
    wg_run_tasks<W, H>(S, I, T, P, lambda, false, false);
    if (tid < GW) {   // pick: the first minimum over the (U, V) task pairs of J_U + J_V + lambda * mode bits
        const int l = tid;
        float J = INF;
        if (2 * l + 1 < S.nTask) {
            const int t = 2 * l, c = S.taskC[t];
            J = S.taskJ[t] + S.taskJ[t + 1];
            if (c == 255) {
                const int aU = S.cflAlpha[0], aV = S.cflAlpha[1];
                const int sU = aU == 0 ? 0 : aU < 0 ? 1 : 2, sV = aV == 0 ? 0 : aV < 0 ? 1 : 2;
                float b = uvbits(UV_CFL_PRED) + C.cfl_sign[sU * 3 + sV - 1];
                if (sU) b += C.cfl_alpha[(sU - 1) * 3 + sV][abs(aU) - 1];
                if (sV) b += C.cfl_alpha[(sV - 1) * 3 + sU][abs(aV) - 1];
                J = fmaf(lambda, b, J);
            } else {
                J = fmaf(lambda, uvcbits(c), J);
            }
        }
        const int w = gargmin(J, l), bt = 2 * (w < 0 ? 0 : w);
        if (l == 0) {
            S.win = bt;
            const int c = S.taskC[bt];
            S.uvMode = c == 255 ? UV_CFL_PRED : cand_mode(c);
            S.uvDelta = c == 255 ? 0 : cand_delta(c);
        }
    }
    __syncthreads();
    const int win = S.win, uvTxType = S.taskT[win], winC = S.taskC[win], nTask0 = S.nTask;
    // 4:4:4 blocks with a 64-sample side hold several chroma tx blocks (nTx > 1): the decision above is the first
    // one's; each further one (raster order, U's then V's) gets its edges from the reconstruction, its own contexts
    // and a closed-loop final quantization with the chosen mode (a one-task run in slot 0 / team 0: the winning pair is
    // (even, odd), so V's kept task survives U's further blocks), and is not coded outside the coded area
    const int nTxX = ss ? 1 : (4 * bw4(bsize)) / W, nTxY = ss ? 1 : (4 * bh4(bsize)) / H;
    for (int p = 1; p <= 2; p++) {
        const int t = win + p - 1;
        int team;
        lds<int16_t> *q = wg_final_levels<W, H>(S, I, T, P, t, team);
        Px *dst = rec_plane<Px>(I, p) + (size_t)cy * I.pitch[p] + cx;
        wg_final_pred<W, H, Px>(S, I, p, t, t, team, dst, I.pitch[p]);
        if (S.taskE[t] > 0) {
            const int fixes = wg_inverse_add<Px>(S, q, txSz, uvTxType, bd, dc_q(I.qindex, bd), ac_q(I.qindex, bd), dst, I.pitch[p]);
            if (fixes && tid == 0) atomicAdd(&B.stats[3], 1u);
        }
        int eob = wg_record_txb(S, B, sb, I, p, txSz, uvTxType, x4, y4, q);
        // only a chroma tx block with a 32-sample side belongs to a 64-sided 4:4:4 block
        if constexpr (W == 32 || H == 32) for (int k = 1; k < nTxX * nTxY; k++) {
            const int ox = (k % nTxX) * W, oy = (k / nTxX) * H;
            if (x4 + (ox >> 2) >= I.MiCols || y4 + (oy >> 2) >= I.MiRows) continue;
            const int r4 = sbr4 + (oy >> 2), c4 = sbc4 + (ox >> 2);
            if (gid() == 0)
                rc_edges<Px>(S, I, p, cx + ox, cy + oy, LW, LH, availLC || ox > 0, availUC || oy > 0, S.bdm[p][r4][c4 + (W >> 2) + 1],
                             S.bdm[p][r4 + (H >> 2) + 1][c4], mi_row, mi_col, availUC, availLC);
            if (gid() == 1) grp_txb_ctx(I, p, txSz, x4 + (ox >> 2), y4 + (oy >> 2), bsize, tx0, ty0, &S.skipCtx[p], &S.dcCtx[p]);
            {
                lds<uint16_t> *sd = S.srcb + (p == 2 ? 2048 : 0);
                const Px *g = src_plane<Px>(I, p) + (size_t)(cy + oy) * I.spitch[p] + cx + ox;
                for (int i = tid; i < W * H; i += RC_NT) sd[i] = g[(size_t)(i >> LW) * I.spitch[p] + (i & (W - 1))];
            }
            if (tid == 0) { S.taskC[0] = winC; S.taskT[0] = uvTxType; S.taskP[0] = p; S.nTask = 1; }
            __syncthreads();
            wg_run_tasks<W, H>(S, I, T, P, lambda, false, false);
            int tm;
            lds<int16_t> *qk = wg_final_levels<W, H>(S, I, T, P, 0, tm);
            Px *dk = rec_plane<Px>(I, p) + (size_t)(cy + oy) * I.pitch[p] + cx + ox;
            wg_final_pred<W, H, Px>(S, I, p, 0, 0, tm, dk, I.pitch[p]);
            if (S.taskE[0] > 0) {
                const int fixes = wg_inverse_add<Px>(S, qk, txSz, uvTxType, bd, dc_q(I.qindex, bd), ac_q(I.qindex, bd), dk, I.pitch[p]);
                if (fixes && tid == 0) atomicAdd(&B.stats[3], 1u);
            }
            eob |= wg_record_txb(S, B, sb, I, p, txSz, uvTxType, x4 + (ox >> 2), y4 + (oy >> 2), qk);
        }
        if (nTxX * nTxY > 1) {   // V's first tx block reads the decision's tasks again
            if (tid == 0) S.nTask = nTask0;
            __syncthreads();
        }
        if (tid == 0) S.winEob[p] = eob;
        __syncthreads();
    }
}

// stage the cost tables of quantizer context qctx (and, when first, the small forward matrices) in LDS
__device__ void stage_tables(lds<RcShared> &S, const DevTables &T, int qctx, bool matrices) {
    const int tid = threadIdx.x;
    if (tid == 0) { S.tab.cost = &T.cost[qctx]; S.tab.iscan = T.iscan; }
    if (matrices) {
        for (int k = tid; k < 3 * 336; k += RC_NT) {
            const int kind = k / 336, r = k % 336, ln = r < 16 ? 2 : r < 80 ? 3 : 4, e = r - (ln == 2 ? 0 : ln == 3 ? 16 : 80);
            S.tab.fwdT[k] = T.fwdT[T.fwdOff[kind][ln] + e];
            S.tab.fwd[k] = T.fwd[T.fwdOff[kind][ln] + e];
        }
        if (tid < 21) S.tab.fwdOff[tid / 7][tid % 7] = (tid / 7) * 336 + ((tid % 7) == 3 ? 16 : (tid % 7) == 4 ? 80 : 0);
        if (tid < 19) S.tab.txScale[tid] = T.txScale[tid];
        for (int k = tid; k < 19 * 16; k += RC_NT) S.tab.txg[k / 16][k % 16] = T.txg[k / 16][k % 16];
        for (int k = tid; k < 19 * 25; k += RC_NT) S.tab.cbOff[k / 25][k % 25] = (&av1t::Coeff_Base_Ctx_Offset[0][0][0])[k];
    }
}

// the closed loop of superblock sb (the workgroup; the tables staged for its image's quantizer context)
template <typename Px>
__device__ void recon_sb(lds<RcShared> &S, const Batch &B, const SearchParams &P, uint32_t sb) {
    const Img &I = B.img[B.sbImg[sb]];
    const DevTables &T = *B.T;
    const uint32_t local = sb - I.sbBase;
    const int sbc = (int)(local % I.sbCols), sbr = (int)(local / I.sbCols), sbx = sbc * 64, sby = sbr * 64;
    const int tid = threadIdx.x;
    int tc = 0, tr = 0;
    while (tc + 1 < I.tileCols && sbx >= I.tileColStart[tc + 1] * 4) tc++;
    while (tr + 1 < I.tileRows && sby >= I.tileRowStart[tr + 1] * 4) tr++;
    const int tx0 = I.tileColStart[tc] * 4, ty0 = I.tileRowStart[tr] * 4;
    const int miColEnd = I.tileColStart[tc + 1], miRowEnd = I.tileRowStart[tr + 1];
    // clear_block_decoded_flags (spec 7.3.3)
    for (int k = tid; k < 3 * 18 * 18; k += RC_NT) {
        const int p = k / 324, yy = (k / 18) % 18 - 1, xx = k % 18 - 1;
        const int ss = p ? I.ss : 0;
        const int sbW4 = (miColEnd - sbc * 16) >> ss, sbH4 = (miRowEnd - sbr * 16) >> ss;
        uint8_t v = (yy < 0 && xx < sbW4) ? 1 : (xx < 0 && yy < sbH4) ? 1 : 0;
        if (yy == (16 >> ss) && xx == -1) v = 0;
        S.bdm[p][yy + 1][xx + 1] = v;
    }
    if (tid == 0) { S.nTxb = 0; S.nCoef = 0; S.flopsOn = B.flops != nullptr; S.flops = 0; }
    __syncthreads();
    const float lambda = T.lambda[I.bd == 10][I.qindex] * P.lambdaMul;
    const int nLeaves = B.leafCount[sb];
    for (int li = 0; li < nLeaves; li++) {
        const Leaf lf = B.leaves[(size_t)sb * MAX_LEAVES + li];
        const int bsize = lf.bsize, x = sbx + lf.x4 * 4, y = sby + lf.y4 * 4, mi_row = y >> 2, mi_col = x >> 2;
        auto inside = [&](int r, int c) { return c >= (tx0 >> 2) && c < miColEnd && r >= (ty0 >> 2) && r < miRowEnd; };
        const int availU = inside(mi_row - 1, mi_col), availL = inside(mi_row, mi_col - 1);
        const int hasC = has_chroma(bsize, mi_row, mi_col, I.mono, I.ss);
        const int availUC = hasC ? ((I.ss && bh4(bsize) == 1) ? inside(mi_row - 2, mi_col) : availU) : 0;
        const int availLC = hasC ? ((I.ss && bw4(bsize) == 1) ? inside(mi_row, mi_col - 2) : availL) : 0;
        const uint32_t firstTxb = (uint32_t)S.nTxb;
        const int txSz = av1t::Max_Tx_Size_Rect[bsize];
#define RC_LUMA(w, h) rc_luma<w, h, Px>(S, B, P, sb, I, bsize, x, y, availU, availL, tx0, ty0, lambda)
        AV1M_TX_SWITCH(txSz, RC_LUMA)
#undef RC_LUMA
        if (hasC) {
            const int uvTx = uv_tx_size(bsize, I.ss);
#define RC_CHROMA(w, h) rc_chroma<w, h, Px>(S, B, P, sb, I, bsize, mi_row, mi_col, availUC, availLC, tx0, ty0, lambda * P.lambdaCMul)
            switch (uvTx) {
            case TX_4X4: RC_CHROMA(4, 4); break; case TX_8X8: RC_CHROMA(8, 8); break; case TX_16X16: RC_CHROMA(16, 16); break;
            case TX_32X32: RC_CHROMA(32, 32); break; case TX_4X8: RC_CHROMA(4, 8); break; case TX_8X4: RC_CHROMA(8, 4); break;
            case TX_8X16: RC_CHROMA(8, 16); break; case TX_16X8: RC_CHROMA(16, 8); break; case TX_16X32: RC_CHROMA(16, 32); break;
            case TX_32X16: RC_CHROMA(32, 16); break; case TX_4X16: RC_CHROMA(4, 16); break; case TX_16X4: RC_CHROMA(16, 4); break;
            case TX_8X32: RC_CHROMA(8, 32); break; default: RC_CHROMA(32, 8); break;
            }
#undef RC_CHROMA
        } else if (tid == 0) {
            S.winEob[1] = S.winEob[2] = 0; S.uvMode = 0; S.uvDelta = 0; S.cflAlpha[0] = S.cflAlpha[1] = 0;
        }
        __syncthreads();
        const int skip = S.winEob[0] == 0 && S.winEob[1] == 0 && S.winEob[2] == 0;
        const int yMode = S.mode, uv = S.uvMode, bw = bw4(bsize), bh = bh4(bsize);
        for (int k = tid; k < bw * bh; k += RC_NT) {
            const int r = mi_row + k / bw, c = mi_col + k % bw;
            if (r < I.MiRows && c < I.MiCols) {
                const size_t o = (size_t)r * I.MiCols + c;
                I.mY[o] = (uint8_t)yMode;
                if (hasC) I.mUV[o] = (uint8_t)uv;
                I.mSkip[o] = (uint8_t)skip;
                I.mMiSize[o] = (uint8_t)bsize;
                I.mTx[o] = (uint8_t)txSz;
            }
        }
        if (tid == 0) {
            RecBlock b;
            b.mi_row = (uint16_t)mi_row; b.mi_col = (uint16_t)mi_col; b.bsize = (uint8_t)bsize; b.skip = (uint8_t)skip; b.tx_depth = 0;
            b.y_mode = (uint8_t)yMode;
            b.angle_delta_y = (int8_t)(bsize >= BLOCK_8X8 && is_dir(yMode) ? S.delta : 0);
            b.uv_mode = (uint8_t)(hasC ? uv : 0);
            b.angle_delta_uv = (int8_t)(hasC && bsize >= BLOCK_8X8 && is_dir(uv) ? S.uvDelta : 0);
            const bool cfl = hasC && uv == UV_CFL_PRED;
            b.cfl_alpha_u = (int8_t)(cfl ? S.cflAlpha[0] : 0); b.cfl_alpha_v = (int8_t)(cfl ? S.cflAlpha[1] : 0);
            b.filter_intra = 0; b.cdef_idx = 0; b.sb_qindex = 0;
            B.oBlocks[(size_t)sb * MAX_LEAVES + li] = b;
            B.oBlockTxb[(size_t)sb * MAX_LEAVES + li] = (uint16_t)firstTxb;
            atomicAdd(&B.stats[0], 1u);
            if (skip) atomicAdd(&B.stats[1], 1u);
            if (cfl) atomicAdd(&B.stats[2], 1u);
        }
        __syncthreads();
    }
    if (tid == 0) {
        SbOut o;
        o.nBlocks = (uint32_t)nLeaves; o.nTxb = (uint32_t)S.nTxb; o.nCoeffs = (uint32_t)S.nCoef; o.pad = 0;
        B.sbOut[sb] = o;
        if (S.flopsOn && S.flops) atomicAdd(&B.flops[32 + (sb & 31)], S.flops);
    }
}

// RcShared allows two workgroups per compute unit.  WPE 2 holds the registers to 256 so both fit (more superblocks in
// flight: batches); WPE 1 is one workgroup per CU, which spends the registers and the rest of the LDS on the latency of
// one superblock (fewer spills, eight transform teams: a single image, whose closed loop is bound by its chain of
// superblocks).  The same results either way.
#define AV1M_RC_KERNEL(WPE) __global__ void __launch_bounds__(RC_NT) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))

// The kernels' one parameter, read in place in the kernarg segment (it is the segment's first argument): handing a
// reference to a by-value kernel parameter to the out-of-line superblock code would make the compiler copy it into
// every lane's scratch, and every read of a field a per-lane scratch load.
struct RcStepArgs {
    Batch B;
    SearchParams P;
    const uint32_t *sbList;
};
struct RcQueueArgs {
    Batch B;
    SearchParams P;
    RcQueue Q;
};
template <typename A> __device__ __forceinline__ const A &kernel_args() { return *(const A *)__builtin_amdgcn_kernarg_segment_ptr(); }

// one launch per wavefront step: workgroup b takes superblock sbList[b]
template <typename Px>
AV1M_RC_KERNEL(2) k_recon(RcStepArgs) {
    extern __shared__ __align__(16) char smraw[];
    lds<RcShared> &S = *(lds<RcShared> *)(lds<char> *)smraw;
    const RcStepArgs &A = kernel_args<RcStepArgs>();
    const Batch &B = A.B;
    const SearchParams &P = A.P;
    const uint32_t sb = A.sbList[blockIdx.x];
    if (threadIdx.x == 0) S.nTeam = RC_NTEAM;
    stage_tables(S, *B.T, B.img[B.sbImg[sb]].qctx, true);
    __syncthreads();
    recon_sb<Px>(S, B, P, sb);
}

// wait (thread 0) until superblock flag f is set; false when the wait ran out of time or another workgroup aborted
__device__ bool wait_done(const uint32_t *f, const RcQueue &Q) {
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (__hip_atomic_load(Q.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if (wall_clock64() - t0 > Q.spinTicks) {
            __hip_atomic_store(Q.abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// Persistent schedule: every workgroup pulls superblocks from the queue (the anti-diagonal wavefront order of all
// images of the batch) and starts each as soon as its left and above-right (above, at a tile's right edge)
// neighbours inside its tile are done.  Items are taken in an order where every dependency was taken earlier by a
// running workgroup, so the waits always end; a wait that exceeds the time limit raises the abort flag, which makes
// every workgroup stop taking items and return.
template <typename Px, int WPE>
AV1M_RC_KERNEL(WPE) k_recon_queue(RcQueueArgs) {
    extern __shared__ __align__(16) char smraw[];
    lds<RcShared> &S = *(lds<RcShared> *)(lds<char> *)smraw;
    const RcQueueArgs &A = kernel_args<RcQueueArgs>();
    const Batch &B = A.B;
    const SearchParams &P = A.P;
    const RcQueue &Q = A.Q;
    const int tid = threadIdx.x;
    int staged = -1;
    if (tid == 0) S.nTeam = WPE == 1 ? 2 * RC_NTEAM : RC_NTEAM;
    for (;;) {
        if (tid == 0) {
            uint32_t it = Q.n;
            if (!__hip_atomic_load(Q.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                it = __hip_atomic_fetch_add(Q.head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t sb = 0xFFFFFFFFu;
            if (it < Q.n) {
                sb = Q.order[it];
                const Img &I = B.img[B.sbImg[sb]];
                const uint32_t local = sb - I.sbBase;
                const int c = (int)(local % I.sbCols), r = (int)(local / I.sbCols);
                int tc = 0, tr = 0;
                while (tc + 1 < I.tileCols && c * 16 >= I.tileColStart[tc + 1]) tc++;
                while (tr + 1 < I.tileRows && r * 16 >= I.tileRowStart[tr + 1]) tr++;
                const int c0 = I.tileColStart[tc] >> 4, c1 = (I.tileColStart[tc + 1] + 15) >> 4, r0 = I.tileRowStart[tr] >> 4;
                bool ok = true;
                if (c > c0) ok = wait_done(Q.done + sb - 1, Q);
                if (ok && r > r0) ok = wait_done(Q.done + sb - I.sbCols + (c + 1 < c1 ? 1 : 0), Q);
                if (!ok) sb = 0xFFFFFFFFu;
            }
            S.qItem = sb;
        }
        __syncthreads();
        const uint32_t sb = S.qItem;
        if (sb == 0xFFFFFFFFu) break;
        // the neighbours' reconstruction, maps and contexts were published with release at agent scope
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const int qctx = B.img[B.sbImg[sb]].qctx;
        if (qctx != staged) {
            stage_tables(S, *B.T, qctx, staged < 0);
            staged = qctx;
        }
        __syncthreads();
        recon_sb<Px>(S, B, P, sb);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) __hip_atomic_store(Q.done + sb, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

static_assert(sizeof(RcShared) <= 80 * 1024, "the closed loop's LDS exceeds half a compute unit's 160 KiB");
static_assert(alignof(RcShared) >= alignof(RcTeam) && sizeof(RcShared) + RC_NTEAM * sizeof(RcTeam) <= 160 * 1024,
              "the eight-team instance's LDS exceeds a compute unit's");
// dynamic LDS of the instance at WPE workgroups per CU
constexpr size_t rc_smem(int wpe) { return sizeof(RcShared) + (wpe == 1 ? RC_NTEAM * sizeof(RcTeam) : 0); }
size_t recon_smem_bytes() { return rc_smem(1); }

int recon_prepare(int bd) {
    const int s2 = (int)rc_smem(2), s1 = (int)rc_smem(1);
    const hipError_t e = bd == 10 ? hipFuncSetAttribute((const void *)k_recon<uint16_t>, hipFuncAttributeMaxDynamicSharedMemorySize, s2)
                                  : hipFuncSetAttribute((const void *)k_recon<uint8_t>, hipFuncAttributeMaxDynamicSharedMemorySize, s2);
    hipError_t f = bd == 10 ? hipFuncSetAttribute((const void *)k_recon_queue<uint16_t, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, s1)
                            : hipFuncSetAttribute((const void *)k_recon_queue<uint8_t, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, s1);
    if (f == hipSuccess)
        f = bd == 10 ? hipFuncSetAttribute((const void *)k_recon_queue<uint16_t, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, s2)
                     : hipFuncSetAttribute((const void *)k_recon_queue<uint8_t, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, s2);
    return e == hipSuccess && f == hipSuccess ? 0 : -1;
}

void launch_recon(const Batch &B, const SearchParams &P, const uint32_t *dSbList, uint32_t n, int bd, hipStream_t st) {
    if (!n) return;
    const RcStepArgs a{B, P, dSbList};
    if (bd == 10) hipLaunchKernelGGL(k_recon<uint16_t>, dim3(n), dim3(RC_NT), rc_smem(2), st, a);
    else hipLaunchKernelGGL(k_recon<uint8_t>, dim3(n), dim3(RC_NT), rc_smem(2), st, a);
}

// resident workgroups of the persistent kernel per compute unit
int recon_queue_slots(int bd, int wpe) {
    int n = 0;
    hipError_t e;
    if (wpe == 2)
        e = bd == 10 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_recon_queue<uint16_t, 2>, RC_NT, rc_smem(2))
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_recon_queue<uint8_t, 2>, RC_NT, rc_smem(2));
    else
        e = bd == 10 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_recon_queue<uint16_t, 1>, RC_NT, rc_smem(1))
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_recon_queue<uint8_t, 1>, RC_NT, rc_smem(1));
    return e == hipSuccess ? n : 0;
}

void launch_recon_queue(const Batch &B, const SearchParams &P, const RcQueue &Q, uint32_t blocks, int bd, int wpe, hipStream_t st) {
    if (!Q.n || !blocks) return;
    const RcQueueArgs a{B, P, Q};
    if (wpe == 2) {
        if (bd == 10) hipLaunchKernelGGL((k_recon_queue<uint16_t, 2>), dim3(blocks), dim3(RC_NT), rc_smem(2), st, a);
        else hipLaunchKernelGGL((k_recon_queue<uint8_t, 2>), dim3(blocks), dim3(RC_NT), rc_smem(2), st, a);
    } else {
        if (bd == 10) hipLaunchKernelGGL((k_recon_queue<uint16_t, 1>), dim3(blocks), dim3(RC_NT), rc_smem(1), st, a);
        else hipLaunchKernelGGL((k_recon_queue<uint8_t, 1>), dim3(blocks), dim3(RC_NT), rc_smem(1), st, a);
    }
}

}  // namespace av1m
