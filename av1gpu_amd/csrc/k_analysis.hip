// Open-loop luma analysis and per-superblock partition DP.
//
// k_analysis<W, H>: 32-lane groups, NG per workgroup, each taking the (image, block position) items of one shape in
// turn (grid stride, the grid capped at a multiple of what the device holds).  Per block:
// edges from the SOURCE with Z-order availability, SATD screening of the 13 modes (integer totals through LDS atomics:
// order free), angle deltas +-1..3 around the two best directional modes, top-K by SATD + lambda_s * mode bits, exact
// RD (forward transform, quantizer, coefficient rate) of the top-K with DCT_DCT, then the other tx types of the set for
// the winner.  Output: BlockRes {J, candidates best first, tx type}.  Reference: encoder/src/k_analysis.cuh:496-689.
//
// k_dp: one wave per 64x64 superblock: bottom-up DP over the 85 square nodes and 10 partition types with the averaged
// partition bits, a level's nodes across the lanes, then the leaves in decode order by prefix sums.  Reference:
// encoder/src/k_dp.cuh:72-297.
#include <type_traits>
#include "dev_eval.h"

namespace av1m {

// Shapes at least 32 wide try DCT_DCT only: their evaluation takes the forward DCT with the residual made on the fly
// from the prediction (grp_fwd_dct: wave-uniform matrix operands), so no residual array sits in LDS.  The levels
// overlay the coefficients (lev): the quantizer's round t writes levels 32 t .. 32 t + 31, i.e. the bytes of
// coefficients 16 t .. 16 t + 15 of the row-major tmp, whose raster index is at most 16 t + 15 -- read in an earlier
// round, or in this one before its store (each lane's store depends on its own load); nothing reads the coefficients
// after the quantizer.  64 x 64 and 64 x 32: 19.7 KB per two-group workgroup, 8 per CU (2 waves per SIMD, the VGPR
// limit) instead of 6.
template <int W, int H>
constexpr bool an_direct() { return W >= 32; }
template <int W, int H>
struct EvalScrDirect {
    static constexpr int TS = W + 1, TH = H > 32 ? 32 : H;
    float tmp[TH * TS];
    static_assert((W > 32 ? 32 : W) * TH * 2 <= TH * TS * 4, "the levels fit in the coefficients' space");
    __device__ int16_t *lev() { return reinterpret_cast<int16_t *>(tmp); }
};
template <int W, int H>
struct AnScr {
    std::conditional_t<an_direct<W, H>(), EvalScrDirect<W, H>, EvalScr<W, H>> ev;
    uint16_t ea[EDGE_LEN], el[EDGE_LEN];   // source edges
    uint16_t pa[EDGE_LEN], pl[EDGE_LEN];   // edge-filtered copies of the candidate under evaluation
    int satdI[32];
    int cand[32];
};

constexpr int an_groups(int W, int H) { return W * H <= 256 ? 8 : W * H <= 1024 ? 4 : 2; }

struct TileBox { int x0, y0, x1, y1; };
__device__ __forceinline__ TileBox tile_box(const Img &I, int x, int y) {
    int c = 0, r = 0;
    while (c + 1 < I.tileCols && x >= I.tileColStart[c + 1] * 4) c++;
    while (r + 1 < I.tileRows && y >= I.tileRowStart[r + 1] * 4) r++;
    return TileBox{I.tileColStart[c] * 4, I.tileRowStart[r] * 4, I.tileColStart[c + 1] * 4, I.tileRowStart[r + 1] * 4};
}
__device__ __forceinline__ int zorder16(int x, int y) {
    int z = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) z |= (((x >> b) & 1) << (2 * b)) | (((y >> b) & 1) << (2 * b + 1));
    return z;
}
// open-loop guess that luma sample (px, py) is decoded before the block at (x, y): the SB row above and the SB to the
// left as the spec's BlockDecoded initialisation, Z-order inside the SB (dev_common.cuh:151-163)
__device__ __forceinline__ int ol_decoded(int px, int py, int x, int y, const TileBox &t) {
    if (px < t.x0 || py < t.y0 || px >= t.x1 || py >= t.y1) return 0;
    const int sbx = x & ~63, sby = y & ~63;
    if (py < sby) return 1;
    if (px < sbx) return py < sby + 64;
    if (px >= sbx + 64 || py >= sby + 64) return 0;
    return zorder16(((px - sbx) >> 2) & 15, ((py - sby) >> 2) & 15) < zorder16(((x - sbx) >> 2) & 15, ((y - sby) >> 2) & 15);
}

// Occupancy target per shape, in waves per SIMD.  At 4 the compiler keeps a kernel under 128 VGPRs and the extra waves
// hide the LDS and table latency, which outweighs the spills that costs most small shapes; the shapes listed at 3
// spill 150-340 bytes per lane at 4 or 5 and run faster with 168 VGPRs (measured on MI355X, kernel trace of the
// 256-image batch with every shape at each of the targets 2 .. 6, the fastest kept per shape: analysis 642 -> 626 ms;
// 8x16 63 -> 56 ms).  The 64-wide shapes take 2: 256 VGPRs and no AGPRs (the two share one 512-entry file per lane,
// so any AGPR beyond 256 VGPRs leaves one wave per SIMD); with the levels over the coefficients that gave 64x64
// 94 -> 49 ms and 64x32 71 -> 56 ms on the same batch.  Where LDS allows fewer waves than the target the compiler
// notes that it cannot reach it; the build silences that.
constexpr int an_wpe(int W, int H) {
    return W == 64 ? 2
         : (W == 8 && H == 16) || (W == 16 && H == 8) || (W == 16 && H == 16) || (W == 16 && H == 32) || (W == 32 && H == 32) ||
                   (W == 32 && H == 64)
             ? 3
             : 4;
}
// HOOKS: the instance that reads an external screener's candidates and / or a block list (AnHook); the default instance
// has neither, so their registers do not weigh on it.
template <int W, int H, int BS, typename Px, bool HOOKS>
__global__ void __launch_bounds__(an_groups(W, H) * 32) __attribute__((amdgpu_waves_per_eu(an_wpe(W, H)))) k_analysis(Batch B, SearchParams P,
                                                                                                      uint32_t nTotal, AnHook hk) {
    constexpr int NG = an_groups(W, H), TXS = tx_of(W, H);
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 4 ? 2 : H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    constexpr int NSUB = (W / 4) * (H / 4), NSX = W / 4;
    __shared__ AnScr<W, H> sm[NG];
    AnScr<W, H> &S = sm[gid()];
    const DevTables &T = *B.T;
    const int lane = glane();
    const uint32_t *starts = B.shStart + (size_t)BS * (B.nimg + 1);
    const int set = tx_set_intra(TXS);
    // a block list (k_an_blank has set every record of the shape to "not analysed"): the groups take its entries
    const uint32_t nItems = (HOOKS && hk.list) ? min(*hk.count, nTotal) : nTotal;
    for (uint32_t item = blockIdx.x * NG + gid(); item < nItems; item += gridDim.x * NG) {
        const uint32_t gb = (HOOKS && hk.list) ? hk.list[item] : item;
        if (HOOKS && gb >= nTotal) continue;   // an entry outside the shape
        int lo = 0, hi = B.nimg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (starts[mid] <= gb) lo = mid; else hi = mid - 1;
        }
        const Img &I = B.img[lo];
        const uint32_t li = gb - starts[lo];
        const int nbx = I.shNbx[BS];
        const int x = (int)(li % nbx) * W, y = (int)(li / nbx) * H;
        BlockRes &out = B.blk[I.shOff[BS] + li];
        if (x >= I.MiCols * 4 || y >= I.MiRows * 4) {
            if (lane == 0) { out.J = INF; out.ncand = 0; }
            continue;
        }
        // learned pruning: a block belongs to the node of its larger side (one of that node's own shapes, or of its
        // 4-way shapes when the sides are 4:1); it is analysed when that node's parent was expanded (the 64x64 node has
        // no parent) and the node keeps that kind of shape (4x4 blocks: the 8x8 node's children only); the screened
        // luma modes are the superblock's shortlist
        uint32_t modeMask = 0x1FFFu;
        if (B.lnMask) {
            const uint32_t *mk = B.lnMask + (size_t)(I.sbBase + (uint32_t)(y >> 6) * I.sbCols + (uint32_t)(x >> 6)) * 12;
            constexpr int M = W > H ? W : H;
            bool keep = true;
            if constexpr (M < 64) {
                constexpr int PN = 2 * M, LV = PN == 64 ? 0 : PN == 32 ? 1 : PN == 16 ? 2 : 3;
                constexpr int BASE = LV == 0 ? 0 : LV == 1 ? 1 : LV == 2 ? 5 : 21;
                const int node = BASE + ((y & 63) / PN) * (64 / PN) + (x & 63) / PN;
                keep = (mk[node >> 5] >> (node & 31)) & 1;
            }
            if constexpr (M > 4) {
                constexpr int LV = M == 64 ? 0 : M == 32 ? 1 : M == 16 ? 2 : 3, BASE = LV == 0 ? 0 : LV == 1 ? 1 : LV == 2 ? 5 : 21;
                constexpr int WORD = (W == 4 * H || H == 4 * W) ? 8 : 4;   // 4-way shapes, or the node's other own shapes
                const int node = BASE + ((y & 63) / M) * (64 / M) + (x & 63) / M;
                keep = keep && ((mk[WORD + (node >> 5)] >> (node & 31)) & 1);
            }
            if (!keep) {
                if (lane == 0) { out.J = INF; out.ncand = 0; }
                continue;
            }
            modeMask = mk[3] & 0x1FFFu;
        }
        const int bd = I.bd;
        const TileBox tb = tile_box(I, x, y);
        const int haveL = x > tb.x0, haveA = y > tb.y0;
        const int haveAR = ol_decoded(x + W, y - 1, x, y, tb), haveBL = ol_decoded(x - 1, y + H, x, y, tb);
        const int maxX = I.MiCols * 4 - 1, maxY = I.MiRows * 4 - 1;
        const int sp = I.spitch[0];
        const Px *plane = (const Px *)I.src[0];
        const Px *src = plane + (size_t)y * sp + x;
        const float lambda = T.lambda[bd == 10][I.qindex] * P.lambdaMul, lamS = P.satdLambda * T.lamS[bd == 10][I.qindex];
        const CostTab &C = T.cost[I.qctx];
        uint16_t *A = S.ea + EDGE_OFF, *L = S.el + EDGE_OFF;
        grp_build_edges(plane, sp, x, y, LW, LH, haveL, haveA, haveAR, haveBL, maxX, maxY, A, L, bd);
        const int dcv = grp_dc_value(A, L, haveL, haveA, LW, LH, bd);
        auto mbits = [&](int c) -> float {
            const int m = cand_mode(c);
            float b = T.avgYMode[m];
            if (BS >= BLOCK_8X8 && is_dir(m)) b += C.angle_delta[m - V_PRED][cand_delta(c) + 3];
            return b;
        };
        const int K = min(P.K, MAXK);
        int top[MAXK];
        int nK = 0;
        // an external screener's candidates, best first, replace the screening (k_analysis.cuh:586-600): entries with a
        // mode of 13 or more are skipped, deltas kept for directional modes of blocks from 8x8 (outside -3..3: 0); a block
        // with no usable entry is screened as without the hook
        if (HOOKS && hk.ext) {
            const ExtCand *ec = hk.ext + (size_t)(hk.extOff[lo] + li) * hk.extTopK;
            const ExtCand c = lane < hk.extTopK ? ec[lane] : ExtCand{0.f, 255, 0, 0, 0};
            const bool ok = c.mode < 13;
            const uint32_t okm = (uint32_t)(__ballot(ok) >> (threadIdx.x & 32));   // this group's half of the wave
            const int rank = __popc(okm & ((1u << lane) - 1));
            if (ok && rank < MAXK) {
                const int d = (BS >= BLOCK_8X8 && is_dir(c.mode) && c.delta >= -3 && c.delta <= 3) ? c.delta : 0;
                S.cand[rank] = cand_pack(c.mode, d);
            }
            nK = min(min(__popc(okm), MAXK), K);
#pragma unroll
            for (int k = 0; k < MAXK; k++) top[k] = k;
            gsync();
        }
        if (!nK) {
        // ---- SATD screening (unfiltered edges) ----
        S.satdI[lane] = 0;
        // the screened modes in mode order (all 13 without a shortlist)
        const int n13 = __popc(modeMask);
        if (lane < 13 && ((modeMask >> lane) & 1)) S.cand[__popc(modeMask & ((1u << lane) - 1))] = cand_pack(lane, 0);
        gsync();
        // blocks of up to 8 4x4s: item = (candidate, pixel), 16 lanes per 4x4 and the Hadamard across them, so a
        // round holds at most two candidates (modes); larger blocks: item = (candidate, 4x4).  A candidate's items are
        // consecutive, so each segment of up to 32 lanes sums its Hadamard costs with shuffles before one LDS atomic.
        auto screen = [&](int c0, int c1) {
            constexpr int NSY = NSUB / NSX;
            if constexpr (NSUB >= 4 && NSUB <= 8 && NSY >= 2) {
                // one candidate per round: lane l sits at (row, column) = ((l >> 2) & 3, l & 3) of every 4x4 its half
                // (l >> 4) covers -- 4x4 columns h * NSX / 2 .. with all their 4x4 rows (NSX >= 2), or the 4x4 rows h,
                // h + 2, .. of the one column -- predicted as vertical runs of pred_col; each 4x4's Hadamard runs
                // across the half's 16 lanes, and the group sums the candidate's costs once
                constexpr int CPH = NSX >= 2 ? NSX / 2 : 1, RUN = NSX >= 2 ? NSY : NSY / 2, DI = NSX >= 2 ? 4 : 8;
                const int h = lane >> 4, r = (lane >> 2) & 3, cc = lane & 3, sy0 = (NSX >= 2 ? 0 : 4 * h) + r;
                int sv[CPH][RUN];   // the lane's source samples, the same for every candidate
#pragma unroll
                for (int k = 0; k < CPH; k++)
#pragma unroll
                    for (int q = 0; q < RUN; q++) sv[k][q] = src[(sy0 + q * DI) * sp + (NSX >= 2 ? h * CPH + k : 0) * 4 + cc];
                for (int ci = c0; ci < c1; ci++) {
                    const int c = S.cand[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
                    int v = 0;
#pragma unroll
                    for (int k = 0; k < CPH; k++) {
                        const int sx = (NSX >= 2 ? h * CPH + k : 0) * 4 + cc;
                        int pr[RUN];
                        pred_col(m, pa, dcv, sy0, DI, sx, LW, LH, A, L, 0, 0, pr);
#pragma unroll
                        for (int q = 0; q < RUN; q++) v += hadamard4_lanes(sv[k][q] - pr[q]);
                    }
                    v = seg_sumi<GW>(v);
                    if (lane == 0) S.satdI[ci] += v;
                }
                gsync();
                return;
            }
            if constexpr (NSUB <= 8) {
                constexpr int PIX = NSUB * 16, SEG = PIX < GW ? PIX : GW, NPL = PIX > GW ? PIX / GW : 1;
                // the pixel of item base + lane is (lane % SEG) + 32 k, k = (base / 32) % NPL, for every candidate: the
                // lane's NPL source samples are loaded once
                int sv[NPL];
#pragma unroll
                for (int k = 0; k < NPL; k++) {
                    const int px = lane % SEG + GW * k, sb = px >> 4;
                    sv[k] = src[((sb / NSX) * 4 + ((px >> 2) & 3)) * sp + (sb % NSX) * 4 + (px & 3)];
                }
                const int items = (c1 - c0) * PIX;
                for (int base = 0; base < items; base += GW) {
                    const int it = base + lane;
                    int v = 0, ci = c0;
                    if (it < items) {
                        ci = c0 + it / PIX;
                        const int px = it % PIX, sb = px >> 4;
                        const int sx = (sb % NSX) * 4 + (px & 3), sy = (sb / NSX) * 4 + ((px >> 2) & 3);
                        const int c = S.cand[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
                        const int p = m == DC_PRED ? dcv : pred_px(m, pa, sy, sx, LW, LH, A, L, 0, 0);
                        int sval = sv[0];
#pragma unroll
                        for (int k = 1; k < NPL; k++) if ((base / GW) % NPL == k) sval = sv[k];
                        v = sval - p;
                    }
                    v = seg_sumi<SEG>(hadamard4_lanes(v));
                    if (it < items && (it & (SEG - 1)) == 0) atomicAdd(&S.satdI[ci], v);
                }
                gsync();
                return;
            }
            if constexpr (NSUB >= GW) {
                // every candidate gives lane l the same 4x4s (l, l + 32, ...): each is loaded once and screened for all
                // the candidates in turn, the group summing each candidate's costs (integers: any order)
                for (int sb = lane; sb < NSUB; sb += GW) {
                    const int sx = (sb % NSX) * 4, sy = (sb / NSX) * 4;
                    int sv[16];
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++) sv[i * 4 + j] = src[(sy + i) * sp + sx + j];
                    for (int ci = c0; ci < c1; ci++) {
                        const int c = S.cand[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
                        int d[16];
#pragma unroll
                        for (int j = 0; j < 4; j++) {   // one run of pred_col per column of the 4x4
                            int r[4];
                            pred_col(m, pa, dcv, sy, 1, sx + j, LW, LH, A, L, 0, 0, r);
#pragma unroll
                            for (int i = 0; i < 4; i++) d[i * 4 + j] = sv[i * 4 + j] - r[i];
                        }
                        const int v = seg_sumi<GW>(hadamard4(d));
                        if (lane == 0) S.satdI[ci] += v;
                    }
                }
                gsync();
                return;
            }
            if constexpr (NSUB == 16) {
                // the same with 16 4x4s: lane l keeps 4x4 l % 16, two candidates a round (one per half)
                const int sb = lane & 15, sx = (sb % NSX) * 4, sy = (sb / NSX) * 4;
                int sv[16];
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) sv[i * 4 + j] = src[(sy + i) * sp + sx + j];
                for (int cq = c0; cq < c1; cq += 2) {
                    const int ci = cq + (lane >> 4);
                    int d[16];
#pragma unroll
                    for (int q = 0; q < 16; q++) d[q] = 0;
                    if (ci < c1) {
                        const int c = S.cand[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            int r[4];
                            pred_col(m, pa, dcv, sy, 1, sx + j, LW, LH, A, L, 0, 0, r);
#pragma unroll
                            for (int i = 0; i < 4; i++) d[i * 4 + j] = sv[i * 4 + j] - r[i];
                        }
                    }
                    const int v = seg_sumi<16>(hadamard4(d));
                    if ((lane & 15) == 0 && ci < c1) S.satdI[ci] += v;
                }
                gsync();
                return;
            }
            constexpr int SEG = NSUB < GW ? NSUB : GW;
            const int items = (c1 - c0) * NSUB;
            for (int base = 0; base < items; base += GW) {
                const int it = base + lane;
                int v = 0, ci = c0;
                if (it < items) {
                    ci = c0 + it / NSUB;
                    const int sb = it % NSUB, sx = (sb % NSX) * 4, sy = (sb / NSX) * 4;
                    const int c = S.cand[ci], m = cand_mode(c), pa = pangle(m, cand_delta(c));
                    int d[16];
#pragma unroll
                    for (int j = 0; j < 4; j++) {   // one run of pred_col per column of the 4x4
                        int r[4];
                        pred_col(m, pa, dcv, sy, 1, sx + j, LW, LH, A, L, 0, 0, r);
#pragma unroll
                        for (int i = 0; i < 4; i++) d[i * 4 + j] = (int)src[(sy + i) * sp + sx + j] - r[i];
                    }
                    v = hadamard4(d);
                }
                v = seg_sumi<SEG>(v);
                if (it < items && (it & (SEG - 1)) == 0) atomicAdd(&S.satdI[ci], v);
            }
            gsync();
        };
        screen(0, n13);
        float js = lane < n13 ? fmaf(lamS, mbits(S.cand[lane]), (float)S.satdI[lane] * 0.25f) : INF;
        if (P.deltas && BS >= BLOCK_8X8) {
            const float jd = (lane < n13 && is_dir(cand_mode(S.cand[lane]))) ? js : INF;
            const int b0 = gargmin(jd, lane);
            const int b1 = gargmin(lane == b0 ? INF : jd, lane);
            const int nd = b0 < 0 ? 0 : b1 < 0 ? 6 : 12;   // deltas around the (up to) two best directional modes
            if (lane < nd) {
                const int k = lane % 6;
                S.cand[13 + lane] = cand_pack(cand_mode(S.cand[lane < 6 ? b0 : b1]), k < 3 ? -(k + 1) : k - 2);
            }
            gsync();
            screen(13, 13 + nd);
            if (lane >= 13 && lane < 13 + nd) js = fmaf(lamS, mbits(S.cand[lane]), (float)S.satdI[lane] * 0.25f);
        }
        for (int k = 0; k < K; k++) {
            top[k] = gargmin(js, lane);
            if (top[k] < 0) break;
            if (lane == top[k]) js = INF;
            nK++;
        }
        }
        // ---- RD of the top-K (DCT_DCT), then the tx types of the set for the winner ----
        const int hasC = has_chroma(BS, y >> 2, x >> 2, I.mono, I.ss);
        const float overhead = (hasC ? P.uvOverheadBits : 0.f) + T.avgSkip0;
        EvalIn e;
        e.txSz = TXS; e.plane = 0; e.dcQ = dc_q(I.qindex, bd); e.acQ = ac_q(I.qindex, bd); e.skipCtx = 0; e.dcCtx = 0;
        e.bd = bd; e.rnd = P.rnd; e.zbin = P.zbin; e.resEnergy = 0.f; e.C = &C; e.intraDir = 0; e.txType = 0;
        float bestJ = INF;
        int bestK = 0, bestTx = DCT_DCT, resK = -1;
        const int nSteps = nK ? nK + ((P.txSearch && set != TX_SET_DCTONLY) ? 15 : 0) : 0;
        for (int st = 0; st < nSteps; st++) {
            const int k = st < nK ? st : bestK, t = st < nK ? DCT_DCT : st - nK + 1;
            if (t != DCT_DCT && !tx_in_set(set, t)) continue;
            const int c = S.cand[top[k]], m = cand_mode(c), pa = pangle(m, cand_delta(c));
            if constexpr (an_direct<W, H>()) {   // DCT_DCT only (st < nK): prediction and forward DCT in one pass
                const uint16_t *EA = A, *EL = L;
                int upA = 0, upL = 0;
                if (I.edgeFilter && is_dir(m)) {
                    for (int i = lane; i < EDGE_LEN; i += GW) { S.pa[i] = S.ea[i]; S.pl[i] = S.el[i]; }
                    gsync();
                    grp_edge_prepare(pa, LW, LH, 0, haveL, haveA, x, y, maxX, maxY, S.pa + EDGE_OFF, S.pl + EDGE_OFF, bd, upA, upL);
                    EA = S.pa + EDGE_OFF;
                    EL = S.pl + EDGE_OFF;
                }
                auto res = [&](int i0, int j, auto &r) {
                    pred_col(m, pa, dcv, i0, 1, j, LW, LH, EA, EL, upA, upL, r);
                    constexpr int CH = sizeof(r) / sizeof(r[0]);
#pragma unroll
                    for (int c = 0; c < CH; c++) r[c] = (int)src[(i0 + c) * sp + j] - r[c];
                };
                lds<float> *tl = (lds<float> *)(lds<char> *)(char *)S.ev.tmp;
                e.resEnergy = grp_fwd_dct<W, H>(tl, T, T.txScale[TXS], res);
                e.txType = DCT_DCT;
                e.intraDir = m;
                const EvalOut o = grp_qrate<W, H>(S.ev.tmp, W + 1, S.ev.lev(), e, T);
                const float J = fmaf(lambda, (o.R + mbits(c)) + overhead, o.D);
                if (J < bestJ) { bestJ = J; bestK = k; }
                gsync();
                continue;
            } else {
            if (k != resK) {
                const uint16_t *EA = A, *EL = L;
                int upA = 0, upL = 0;
                if (I.edgeFilter && is_dir(m)) {
                    for (int i = lane; i < EDGE_LEN; i += GW) { S.pa[i] = S.ea[i]; S.pl[i] = S.el[i]; }
                    gsync();
                    grp_edge_prepare(pa, LW, LH, 0, haveL, haveA, x, y, maxX, maxY, S.pa + EDGE_OFF, S.pl + EDGE_OFF, bd, upA, upL);
                    EA = S.pa + EDGE_OFF;
                    EL = S.pl + EDGE_OFF;
                }
                for (int p = lane; p < W * H; p += GW) {
                    const int i = p >> LW, j = p & (W - 1);
                    const int pr = m == DC_PRED ? dcv : pred_px(m, pa, i, j, LW, LH, EA, EL, upA, upL);
                    S.ev.res[p] = (int16_t)((int)src[i * sp + j] - pr);
                }
                gsync();
                e.resEnergy = (W == 64 || H == 64) ? grp_res_energy<W, H>(S.ev.res) : 0.f;
                resK = k;
            }
            e.txType = t;
            e.intraDir = m;
            const EvalOut o = grp_eval<W, H>(S.ev.res, S.ev.tmp, S.ev.q, e, T);
            const float J = fmaf(lambda, (o.R + mbits(c)) + overhead, o.D);
            if (J < bestJ) {
                bestJ = J;
                if (st < nK) bestK = k; else bestTx = t;
            }
            gsync();
            }
        }
        if (lane == 0) {
            out.J = nK ? bestJ : INF;
            out.tx = (uint8_t)bestTx;
            int n = 0;
            if (nK) out.cand[n++] = (uint8_t)S.cand[top[bestK]];
            for (int k = 0; k < nK; k++)
                if (k != bestK) out.cand[n++] = (uint8_t)S.cand[top[k]];
            out.ncand = (uint8_t)n;
            // work accounting: nK evaluations with DCT_DCT, then the set's other types (the analysis's evaluations)
            constexpr uint32_t FDCT = fwd_flops(W, H, DCT_DCT), FSET = set_flops(W, H, tx_set_intra(TXS));
            if (B.flops && nK) atomicAdd(&B.flops[gb & 31], (unsigned long long)(nK * FDCT + (nSteps > nK ? FSET : 0u)));
        }
        gsync();
    }
}

// every record of a shape "not analysed" (J = INF, no candidate): the blocks a block list leaves out
__global__ void k_an_blank(BlockRes *blk, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        BlockRes r;
        r.J = INF;
        for (int k = 0; k < MAXK; k++) r.cand[k] = 0;
        r.ncand = 0;
        r.tx = 0;
        blk[i] = r;
    }
}

// workgroups of a shape's launch: the kernels loop over the blocks (grid stride), so the grid is capped at
// AN_WAVES_PER_SLOT times what the device holds at once -- most blocks of the small shapes are pruned by the learned
// network and take a few instructions each, which one workgroup per 8 blocks spent on workgroup dispatch instead.
// Measured on MI355X (analysis of the 256-image batch): 908 ms uncapped, 836 / 811 / 771 / 771 / 780 / 798 ms at 2 /
// 4 / 16 / 32 / 64 / 128 (few workgroups balance the unevenly pruned blocks worse)
constexpr int AN_WAVES_PER_SLOT = 16;
// twice that for the shapes that measured faster at 32 (kernel trace of the same batch, every shape at 8 / 16 / 32:
// 4x4 29.2 -> 27.8 ms, 64x32 55.7 -> 54.6, 64x64 49.4 -> 48.6, 4x8, 8x4, 16x32, 32x64 0.5-0.9 ms each)
constexpr int an_slot_mult(int W, int H) {
    return (W == 4 && H <= 8) || (W == 8 && H == 4) || (W == 16 && H == 32) || (W == 32 && H == 64) || (W == 64 && H >= 32) ? 2 : 1;
}
template <typename K>
static uint32_t an_grid_cap(K kernel, int threads, int mult) {
    int dev = 0, cus = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess || cus <= 0 || occ <= 0)
        return 1u << 30;
    return (uint32_t)cus * (uint32_t)occ * AN_WAVES_PER_SLOT * (uint32_t)mult;
}

template <int W, int H, int BS>
static void launch_an(const Batch &B, const SearchParams &P, uint32_t n, int bd, const AnHooks &h, hipStream_t st) {
    if (!n) return;
    constexpr int NG = an_groups(W, H);
    const AnHook hk = h.hook ? h.hook[BS] : AnHook{};
    if (hk.list) hipLaunchKernelGGL(k_an_blank, dim3((n + 255) / 256), dim3(256), 0, st, B.blk + h.base[BS], n);
    // the cap of this instance on the current device (host-side queries, per launch: encoders on several host threads
    // share nothing here)
    const bool hooked = hk.ext || hk.list;
    constexpr int M = an_slot_mult(W, H);
    const uint32_t c = bd == 10 ? (hooked ? an_grid_cap(k_analysis<W, H, BS, uint16_t, true>, NG * 32, M)
                                          : an_grid_cap(k_analysis<W, H, BS, uint16_t, false>, NG * 32, M))
                                : (hooked ? an_grid_cap(k_analysis<W, H, BS, uint8_t, true>, NG * 32, M)
                                          : an_grid_cap(k_analysis<W, H, BS, uint8_t, false>, NG * 32, M));
    uint32_t grid = (n + NG - 1) / NG;
    if (grid > c) grid = c;
    if (bd == 10) {
        if (hooked) hipLaunchKernelGGL((k_analysis<W, H, BS, uint16_t, true>), dim3(grid), dim3(NG * 32), 0, st, B, P, n, hk);
        else hipLaunchKernelGGL((k_analysis<W, H, BS, uint16_t, false>), dim3(grid), dim3(NG * 32), 0, st, B, P, n, hk);
    } else {
        if (hooked) hipLaunchKernelGGL((k_analysis<W, H, BS, uint8_t, true>), dim3(grid), dim3(NG * 32), 0, st, B, P, n, hk);
        else hipLaunchKernelGGL((k_analysis<W, H, BS, uint8_t, false>), dim3(grid), dim3(NG * 32), 0, st, B, P, n, hk);
    }
}

// all shapes of the batch (shTotal[s]: blocks of shape s over the batch; 0 = not analysed), largest first, the i-th
// shape on sts[i % ns] (the shapes are independent: several streams fill each other's launch tails)
void launch_analysis_on(const Batch &B, const SearchParams &P, const uint32_t *shTotal, int bd, const AnHooks &h, const hipStream_t *sts,
                        int ns) {
    int i = 0;
    auto st = [&]() { return sts[(i++) % ns]; };
    launch_an<64, 64, BLOCK_64X64>(B, P, shTotal[BLOCK_64X64], bd, h, st());
    launch_an<64, 32, BLOCK_64X32>(B, P, shTotal[BLOCK_64X32], bd, h, st());
    launch_an<32, 64, BLOCK_32X64>(B, P, shTotal[BLOCK_32X64], bd, h, st());
    launch_an<64, 16, BLOCK_64X16>(B, P, shTotal[BLOCK_64X16], bd, h, st());
    launch_an<16, 64, BLOCK_16X64>(B, P, shTotal[BLOCK_16X64], bd, h, st());
    launch_an<32, 32, BLOCK_32X32>(B, P, shTotal[BLOCK_32X32], bd, h, st());
    launch_an<32, 16, BLOCK_32X16>(B, P, shTotal[BLOCK_32X16], bd, h, st());
    launch_an<16, 32, BLOCK_16X32>(B, P, shTotal[BLOCK_16X32], bd, h, st());
    launch_an<32, 8, BLOCK_32X8>(B, P, shTotal[BLOCK_32X8], bd, h, st());
    launch_an<8, 32, BLOCK_8X32>(B, P, shTotal[BLOCK_8X32], bd, h, st());
    launch_an<16, 16, BLOCK_16X16>(B, P, shTotal[BLOCK_16X16], bd, h, st());
    launch_an<16, 8, BLOCK_16X8>(B, P, shTotal[BLOCK_16X8], bd, h, st());
    launch_an<8, 16, BLOCK_8X16>(B, P, shTotal[BLOCK_8X16], bd, h, st());
    launch_an<16, 4, BLOCK_16X4>(B, P, shTotal[BLOCK_16X4], bd, h, st());
    launch_an<4, 16, BLOCK_4X16>(B, P, shTotal[BLOCK_4X16], bd, h, st());
    launch_an<8, 8, BLOCK_8X8>(B, P, shTotal[BLOCK_8X8], bd, h, st());
    launch_an<8, 4, BLOCK_8X4>(B, P, shTotal[BLOCK_8X4], bd, h, st());
    launch_an<4, 8, BLOCK_4X8>(B, P, shTotal[BLOCK_4X8], bd, h, st());
    launch_an<4, 4, BLOCK_4X4>(B, P, shTotal[BLOCK_4X4], bd, h, st());
}
void launch_analysis(const Batch &B, const SearchParams &P, const uint32_t *shTotal, int bd, const AnHooks &h, hipStream_t st) {
    launch_analysis_on(B, P, shTotal, bd, h, &st, 1);
}

// ---------------------------------------------------------------------------------------------------------------
// Partition DP: one wave per 64x64 superblock, four superblocks per workgroup.  The 85 square nodes are solved level by
// level from the 8x8 nodes up, one node per lane (64, 16, 4, 1 lanes); a node weighs every partition the frame edge and
// the options allow (Part_Order), each costing fmaf(lambda, bits, sum of its in-frame parts' J in part order) -- the
// analysis J of a block part, the already solved cost of a sub-node -- and keeps the first cheapest.  The leaves are
// then placed without a walk: every node knows how many leaves its chosen subtree has, offsets flow down the split
// nodes level by level, and each node that ends a branch writes its own blocks at its offset, so the decode order of
// the reference's walk (k_dp.cuh:72-297) comes out of the prefix sums.
constexpr int DP_WAVES = 4;
constexpr uint16_t DP_DEAD = 0xFFFF;   // a node no chosen tree reaches
struct DpNodes {
    float J[85];
    int8_t part[85];   // -1: the node lies outside the frame
    uint16_t leaves[85], first[85];
};
__device__ __forceinline__ int dp_node(int L, int nx, int ny) { return (L == 0 ? 0 : L == 1 ? 1 : L == 2 ? 5 : 21) + ny * (1 << L) + nx; }

__global__ void __launch_bounds__(64 * DP_WAVES) k_dp(Batch B, SearchParams P) {
    __shared__ DpNodes S[DP_WAVES];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t sb = blockIdx.x * DP_WAVES + w;
    const bool active = sb < B.nSb;
    DpNodes &D = S[w];
    const Img *Ip = active ? &B.img[B.sbImg[sb]] : nullptr;
    int sbx = 0, sby = 0, FW = 0, FH = 0;
    float lambda = 0.f;
    if (active) {
        const uint32_t local = sb - Ip->sbBase;
        sbx = (int)(local % Ip->sbCols) * 64; sby = (int)(local / Ip->sbCols) * 64;
        FW = Ip->MiCols * 4; FH = Ip->MiRows * 4;
        lambda = B.T->lambda[Ip->bd == 10][Ip->qindex] * P.lambdaMul;
    }
    auto blockJ = [&](int bs, int x, int y) -> float {
        if (bs >= BLOCK_INVALID || Ip->shNbx[bs] == 0) return INF;
        const int bw = 4 * bw4(bs), bh = 4 * bh4(bs);
        return B.blk[Ip->shOff[bs] + (uint32_t)(y / bh) * Ip->shNbx[bs] + x / bw].J;
    };
    // bottom-up: node costs, choices, leaf counts
    for (int L = 3; L >= 0; L--) {
        const int per = 1 << L, n = 64 >> L;
        if (active && lane < per * per) {
            const int nx = lane % per, ny = lane / per, ni = dp_node(L, nx, ny);
            const int x = sbx + nx * n, y = sby + ny * n, q = n >> 2;
            if (x >= FW || y >= FH) {
                D.J[ni] = 0.f; D.part[ni] = -1; D.leaves[ni] = 0;
                if (B.dpRel) for (int k = 0; k < 3; k++) B.dpRel[(size_t)sb * 255 + 85 * k + ni] = 0.f;
            } else {
                const bool rows = y + 2 * q < FH, cols = x + 2 * q < FW, big = n > P.maxPart;
                const int pmask = n == 8 ? P.partMask8 : P.partMaskN, bsl = L == 3 ? 1 : L == 2 ? 2 : L == 1 ? 3 : 4;
                // what the frame edge leaves: inside, every partition the options allow at avgPart bits; past the
                // bottom (right) edge SPLIT or HORZ (VERT) for one bit; past both, SPLIT for none
                uint32_t allowed;
                float edgeBits;
                if (rows && cols) { allowed = big ? (1u << PARTITION_SPLIT) : ((uint32_t)pmask | (1u << PARTITION_SPLIT)); edgeBits = -1.f; }
                else if (cols) { allowed = (1u << PARTITION_SPLIT) | (big ? 0u : 1u << PARTITION_HORZ); edgeBits = 1.f; }
                else if (rows) { allowed = (1u << PARTITION_SPLIT) | (big ? 0u : 1u << PARTITION_VERT); edgeBits = 1.f; }
                else { allowed = 1u << PARTITION_SPLIT; edgeBits = 0.f; }
                if (n == 8) allowed &= 0xFu;   // no AB or 4-way partitions of 8x8 nodes
                float best = INF, noKids = INF, noOwn = INF, noFour = INF;
                int bp = PARTITION_SPLIT;
                uint16_t bLeaves = 0;
                for (int o = 0; o < 10; o++) {
                    const int p = Part_Order[o];
                    if (!((allowed >> p) & 1)) continue;
                    float J = 0.f;
                    uint16_t nl = 0;
                    for (int k = 0; k < Part_Shape[p].parts; k++) {
                        const PartPlace at = Part_Shape[p].at[k];
                        const int px = x + at.dx * q, py = y + at.dy * q;
                        if (px >= FW || py >= FH) continue;
                        if (p == PARTITION_SPLIT && n > 8) {
                            const int ci = dp_node(L + 1, nx * 2 + (at.dx >> 1), ny * 2 + (at.dy >> 1));
                            J = J + D.J[ci];
                            nl += D.leaves[ci];
                        } else {
                            J = J + blockJ(part_bsize(p, k, n), px, py);
                            nl++;
                        }
                    }
                    if (p == PARTITION_NONE) {
                        if (!(J < INF)) continue;
                        J = J * P.splitBias;
                    }
                    const float bits = edgeBits < 0.f ? B.T->avgPart[bsl][p] : edgeBits;
                    const float t = fmaf(lambda, bits, J);
                    if (t < best) { best = t; bp = p; bLeaves = nl; }
                    const bool four = p == PARTITION_HORZ_4 || p == PARTITION_VERT_4;
                    if (four || p == PARTITION_NONE || p == PARTITION_HORZ || p == PARTITION_VERT) noKids = fminf(noKids, t);
                    if (four || p == PARTITION_SPLIT) noOwn = fminf(noOwn, t);
                    if (!four) noFour = fminf(noFour, t);
                }
                D.J[ni] = best; D.part[ni] = (int8_t)bp; D.leaves[ni] = bLeaves;
                if (B.dpRel) {   // relative increase of the node's cost without its children / own shapes / 4-way shapes
                    const float den = fmaxf(best, 1e-3f);
                    B.dpRel[(size_t)sb * 255 + ni] = fminf((noKids - best) / den, 1e6f);
                    B.dpRel[(size_t)sb * 255 + 85 + ni] = fminf((noOwn - best) / den, 1e6f);
                    B.dpRel[(size_t)sb * 255 + 170 + ni] = fminf((noFour - best) / den, 1e6f);
                }
            }
            D.first[ni] = ni == 0 ? 0 : DP_DEAD;
        }
        __syncthreads();
    }
    // top-down: each reached split node hands its children their first leaf index (children in decode order)
    for (int L = 0; L < 3; L++) {
        const int per = 1 << L;
        if (active && lane < per * per) {
            const int nx = lane % per, ny = lane / per, ni = dp_node(L, nx, ny);
            if (D.first[ni] != DP_DEAD && D.part[ni] == PARTITION_SPLIT) {
                uint16_t f = D.first[ni];
                for (int k = 0; k < 4; k++) {
                    const int ci = dp_node(L + 1, nx * 2 + (k & 1), ny * 2 + (k >> 1));
                    D.first[ci] = f;
                    f += D.leaves[ci];
                }
            }
        }
        __syncthreads();
    }
    // every reached node that ends a branch writes its blocks
    if (!active) return;
    Leaf *out = B.leaves + (size_t)sb * MAX_LEAVES;
    for (int L = 0; L < 4; L++) {
        const int per = 1 << L, n = 64 >> L;
        if (lane >= per * per) continue;
        const int nx = lane % per, ny = lane / per, ni = dp_node(L, nx, ny), p = D.part[ni];
        if (D.first[ni] == DP_DEAD || p < 0 || (p == PARTITION_SPLIT && n > 8)) continue;
        const int x = sbx + nx * n, y = sby + ny * n, q = n >> 2;
        int at = D.first[ni];
        for (int k = 0; k < Part_Shape[p].parts; k++) {
            const int px = x + Part_Shape[p].at[k].dx * q, py = y + Part_Shape[p].at[k].dy * q;
            if (px >= FW || py >= FH) continue;
            Leaf lf;
            lf.bsize = (uint8_t)part_bsize(p, k, n); lf.x4 = (uint8_t)((px - sbx) >> 2); lf.y4 = (uint8_t)((py - sby) >> 2); lf.pad = 0;
            out[at++] = lf;
        }
    }
    if (lane == 0) { B.sbJ[sb] = D.J[0]; B.leafCount[sb] = D.leaves[0]; }
}

void launch_dp(const Batch &B, const SearchParams &P, hipStream_t st) {
    if (!B.nSb) return;
    hipLaunchKernelGGL(k_dp, dim3((B.nSb + DP_WAVES - 1) / DP_WAVES), dim3(64 * DP_WAVES), 0, st, B, P);
}

}  // namespace av1m
