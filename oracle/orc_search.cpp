// Oracle (test infrastructure only): serial CPU restatement of the reference's GPU intra RD search
// (encoder/src/dev_common.cuh, dev_tx.cuh, rc_eval.cuh, k_analysis.cuh, k_analysis_small.cuh, k_dp.cuh, k_recon.cuh,
// k_misc.cu, pipeline.cpp) for the plain search: every luma mode with angle deltas, chroma modes with deltas and CfL,
// all 19 shapes, all partitions, all intra tx types, no pruning / shortlists / palettes / filter intra / QM / delta-q.
//
// Float semantics (the contract the HIP kernels are checked against bit for bit):
//  * forward transform: each output one fmaf chain over k in ascending order (column pass, then row pass, the row
//    output multiplied by txScale); identity kinds one multiply; 64-point sizes keep the top-left 32 outputs;
//  * quantizer: a = (|v| * den) / step (IEEE division), level = a < zbin ? 0 : (int)(a + rnd), clamp 16383;
//  * sums over coefficients (distortion, 64-point kept energy, rate, residual energy) use 32 virtual lanes: lane l
//    accumulates the elements l, l+32, l+64, ... in order, then the lanes are combined by the xor butterfly
//    16, 8, 4, 2, 1 (dev_common.cuh:77-80);
//  * J = fmaf(lambda, bits, D) wherever the reference forms D + lambda * bits; screening costs fmaf(lamS, bits, satd);
//  * CfL alpha from exact integer sums (the reference's float sums are exact wherever they do not exceed 2^24).
// Tie rules: the first minimum in candidate / task order wins.
#include "orc_search.h"
#include <math.h>
#include <algorithm>
#include <thread>
#include <atomic>

namespace orc {

// ------------------------------------------------------------------------------------------------------------
// tables (k_misc.cu:477-644, dev_common.cuh:99-107)
// ------------------------------------------------------------------------------------------------------------
static std::vector<double> invert(std::vector<double> a, int n) {   // Gauss-Jordan, partial pivoting (k_misc.cu:477-502)
    std::vector<double> inv((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (fabs(a[(size_t)r * n + c]) > fabs(a[(size_t)p * n + c])) p = r;
        if (p != c)
            for (int k = 0; k < n; k++) { std::swap(a[(size_t)c * n + k], a[(size_t)p * n + k]); std::swap(inv[(size_t)c * n + k], inv[(size_t)p * n + k]); }
        const double d = a[(size_t)c * n + c];
        for (int k = 0; k < n; k++) { a[(size_t)c * n + k] /= d; inv[(size_t)c * n + k] /= d; }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; k++) {
                a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
    return inv;
}

// costs of one CDF of nsym symbols (array of nsym + 1 entries, the last being the counter)
static void cdf_cost(const uint16_t *cdf, int nsym, float *out) {
    for (int s = 0; s < nsym; s++) {
        const double lo = s ? cdf[s - 1] : 0.0, hi = cdf[s];
        const double pr = std::max((hi - lo) / 32768.0, 1.0 / 32768.0);
        out[s] = (float)-std::log2(pr);
    }
}
template <typename T, size_t N, typename U> static void cost_arr(const T *cdf, int nsym, int count, U &out) {
    float *o = (float *)&out;
    for (int i = 0; i < count; i++) cdf_cost(cdf + (size_t)i * (nsym + 1), nsym, o + (size_t)i * nsym);
    (void)sizeof(N);
}
#define COSTS(dst, src, nsym) cost_arr<uint16_t, 0>((const uint16_t *)(src), nsym, (int)(sizeof(dst) / sizeof(float) / (nsym)), dst)

static void build_costs(CostTab &C, int q) {
    using namespace av1t;
    COSTS(C.y_mode, Default_Intra_Frame_Y_Mode_Cdf, 13);
    COSTS(C.uv_cfl_no, Default_Uv_Mode_Cfl_Not_Allowed_Cdf, 13);
    COSTS(C.uv_cfl_yes, Default_Uv_Mode_Cfl_Allowed_Cdf, 14);
    COSTS(C.angle_delta, Default_Angle_Delta_Cdf, 7);
    COSTS(C.skip, Default_Skip_Cdf, 2);
    COSTS(C.tx_set1, Default_Intra_Tx_Type_Set1_Cdf, 7);
    COSTS(C.tx_set2, Default_Intra_Tx_Type_Set2_Cdf, 5);
    COSTS(C.cfl_sign, Default_Cfl_Sign_Cdf, 8);
    COSTS(C.cfl_alpha, Default_Cfl_Alpha_Cdf, 16);
    COSTS(C.txb_skip, Default_Txb_Skip_Cdf[q], 2);
    COSTS(C.eob_pt16, Default_Eob_Pt_16_Cdf[q], 5);
    COSTS(C.eob_pt32, Default_Eob_Pt_32_Cdf[q], 6);
    COSTS(C.eob_pt64, Default_Eob_Pt_64_Cdf[q], 7);
    COSTS(C.eob_pt128, Default_Eob_Pt_128_Cdf[q], 8);
    COSTS(C.eob_pt256, Default_Eob_Pt_256_Cdf[q], 9);
    COSTS(C.eob_pt512, Default_Eob_Pt_512_Cdf[q], 10);
    COSTS(C.eob_pt1024, Default_Eob_Pt_1024_Cdf[q], 11);
    COSTS(C.eob_extra, Default_Eob_Extra_Cdf[q], 2);
    COSTS(C.dc_sign, Default_Dc_Sign_Cdf[q], 2);
    COSTS(C.coeff_base_eob, Default_Coeff_Base_Eob_Cdf[q], 3);
    COSTS(C.coeff_base, Default_Coeff_Base_Cdf[q], 4);
    COSTS(C.coeff_br, Default_Coeff_Br_Cdf[q], 4);
    COSTS(C.tx8, Default_Tx_8x8_Cdf, 2);
    COSTS(C.tx16, Default_Tx_16x16_Cdf, 3);
    COSTS(C.tx32, Default_Tx_32x32_Cdf, 3);
    COSTS(C.tx64, Default_Tx_64x64_Cdf, 3);
}

// libaom-style lambda (dev_common.cuh:99-107) in double, rounded to float once
static float lambda_of(int qindex, int bd) {
    const int q = av1t::Dc_Qlookup[(bd - 8) >> 1][qindex];
    double rd = (double)q * q * (3.3 + 0.0015 * q);
    int w = ((255 - qindex) * 3) / 4;
    w = w < 0 ? 0 : w > 72 ? 72 : w;
    rd = rd * (128 + w) / 128.0;
    return (float)(rd / 2048.0 * 1.0);
}

static Tables build_tables() {
    Tables T;
    double norm2[3][7] = {};
    for (int k = 0; k < 3; k++)
        for (int ln = 2; ln <= 6; ln++) {
            if ((k == K_ADST && ln > 4) || (k == K_IDT && ln > 5)) continue;
            const int n = 1 << ln, A = 1 << 14;
            std::vector<double> M((size_t)n * n);
            double nn = 0;
            for (int j = 0; j < n; j++) {
                int32_t t[64] = {0};
                t[j] = A;
                uint32_t e = 0;
                inv_1d(t, k, ln, 31, &e);
                for (int i = 0; i < n; i++) {
                    M[(size_t)i * n + j] = (double)t[i] / A;
                    nn += M[(size_t)i * n + j] * M[(size_t)i * n + j];
                }
            }
            norm2[k][ln] = nn / n;
            const std::vector<double> F = invert(M, n);
            T.fwd[k][ln].resize((size_t)n * n);
            for (size_t i = 0; i < (size_t)n * n; i++) T.fwd[k][ln][i] = (float)F[i];
        }
    for (int t = 0; t < 19; t++) {
        const int lw = av1t::Tx_Width_Log2[t], lh = av1t::Tx_Height_Log2[t];
        const double rect = (lw - lh == 1 || lh - lw == 1) ? 2896.0 / 4096.0 : 1.0;
        const double s = (double)(1 << (av1t::Transform_Row_Shift[t] + 4));
        T.txScale[t] = (float)(s / rect);
        for (int ty = 0; ty < 16; ty++) {
            const double nr = norm2[row_kind(ty)][lw], nc = norm2[col_kind(ty)][lh];
            T.txg[t][ty] = (nr == 0 || nc == 0) ? 0.f : (float)(rect * rect * nr * nc / (s * s));
        }
    }
    for (int q = 0; q < 4; q++) build_costs(T.cost[q], q);
    // open-loop averages from the q-context 0 tables (k_misc.cu:590-616)
    for (int m = 0; m < 13; m++) {
        double p = 0;
        for (int a = 0; a < 5; a++)
            for (int l = 0; l < 5; l++) {
                const uint16_t *cd = av1t::Default_Intra_Frame_Y_Mode_Cdf[a][l];
                p += (cd[m] - (m ? cd[m - 1] : 0)) / 32768.0;
            }
        T.avgYMode[m] = (float)-std::log2(std::max(p / 25.0, 1e-6));
    }
    for (int b = 0; b < 5; b++)
        for (int p = 0; p < 10; p++) T.avgPart[b][p] = 20.f;
    for (int bsl = 1; bsl <= 4; bsl++) {
        const int N = bsl == 1 ? 4 : 10;
        for (int p = 0; p < N; p++) {
            double pr = 0;
            for (int ctx = 0; ctx < 4; ctx++) {
                const uint16_t *cd = bsl == 1 ? av1t::Default_Partition_W8_Cdf[ctx] : bsl == 2 ? av1t::Default_Partition_W16_Cdf[ctx]
                                   : bsl == 3 ? av1t::Default_Partition_W32_Cdf[ctx] : av1t::Default_Partition_W64_Cdf[ctx];
                pr += (cd[p] - (p ? cd[p - 1] : 0)) / 32768.0;
            }
            T.avgPart[bsl][p] = (float)-std::log2(std::max(pr / 4.0, 1e-6));
        }
    }
    double ps = 0;
    for (int ctx = 0; ctx < 3; ctx++) ps += av1t::Default_Skip_Cdf[ctx][0] / 32768.0;
    T.avgSkip0 = (float)-std::log2(ps / 3.0);
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 256; q++) {
            T.lambda[b][q] = lambda_of(q, b ? 10 : 8);
            T.lamS[b][q] = 1.0f * sqrtf(T.lambda[b][q]);
        }
    return T;
}
const Tables &tables() {
    static const Tables T = build_tables();
    return T;
}

// ------------------------------------------------------------------------------------------------------------
// evaluation of one tx block (dev_tx.cuh:47-316, rc_eval.cuh:152-226)
// ------------------------------------------------------------------------------------------------------------
static float lane_sum(float *v) {   // xor butterfly over 32 virtual lanes (dev_common.cuh:77-80)
    for (int o = 16; o > 0; o >>= 1) {
        float n[32];
        for (int l = 0; l < 32; l++) n[l] = v[l] + v[l ^ o];
        memcpy(v, n, sizeof n);
    }
    return v[0];
}

static inline int txsz_ctx(int t) { return (av1t::Tx_Size_Sqr[t] + av1t::Tx_Size_Sqr_Up[t] + 1) >> 1; }

static float base_bits(const EvalIn &e, int eob) {   // dev_tx.cuh:47-86
    const CostTab &C = *e.C;
    const int t = e.txSz, ctx = txsz_ctx(t);
    if (eob == 0) return C.txb_skip[ctx][e.skipCtx][1];
    float b = C.txb_skip[ctx][e.skipCtx][0];
    const int ptype = e.plane > 0;
    if (e.plane == 0) {
        const int set = tx_set_intra(t);
        if (set > 0) {
            const int sqr = av1t::Tx_Size_Sqr[t];
            int idx = 0;
            if (set == TX_SET_INTRA_1) { for (int i = 0; i < 7; i++) if (av1t::Tx_Type_Intra_Inv_Set1[i] == e.txType) idx = i; b += C.tx_set1[sqr][e.intraDir][idx]; }
            else { for (int i = 0; i < 5; i++) if (av1t::Tx_Type_Intra_Inv_Set2[i] == e.txType) idx = i; b += C.tx_set2[sqr][e.intraDir][idx]; }
        }
    }
    const int ms = std::min((int)av1t::Tx_Width_Log2[t], 5) + std::min((int)av1t::Tx_Height_Log2[t], 5) - 4;
    int eobPt;
    if (eob <= 2) eobPt = eob;
    else { eobPt = 2; int v = eob - 1; while (v > 1) { v >>= 1; eobPt++; } }
    const int ectx = tx_class(e.txType) == TX_CLASS_2D ? 0 : 1;
    float pt;
    switch (ms) {
    case 0: pt = C.eob_pt16[ptype][ectx][eobPt - 1]; break;
    case 1: pt = C.eob_pt32[ptype][ectx][eobPt - 1]; break;
    case 2: pt = C.eob_pt64[ptype][ectx][eobPt - 1]; break;
    case 3: pt = C.eob_pt128[ptype][ectx][eobPt - 1]; break;
    case 4: pt = C.eob_pt256[ptype][ectx][eobPt - 1]; break;
    case 5: pt = C.eob_pt512[ptype][eobPt - 1]; break;
    default: pt = C.eob_pt1024[ptype][eobPt - 1]; break;
    }
    b += pt;
    if (eobPt >= 3) {
        const int rem = eob - ((1 << (eobPt - 2)) + 1);
        b += C.eob_extra[ctx][ptype][eobPt - 3][(rem >> (eobPt - 3)) & 1];
        b += (float)(eobPt - 3);
    }
    return b;
}

static float coeff_bits(int c, int eob, const int16_t *q, const uint16_t *scan, int txSz, int txClass, int ptype, int dcCtx,
                        const CostTab &C) {   // dev_tx.cuh:88-165
    const int ctx = txsz_ctx(txSz), adj = av1t::Adjusted_Tx_Size[txSz];
    const int bwl = av1t::Tx_Width_Log2[adj], width = 1 << bwl, height = av1t::Tx_Height[adj];
    const int pos = scan[c], qv = q[pos], v = qv < 0 ? -qv : qv, level = v < 15 ? v : 15;
    const int row = pos >> bwl, col = pos - (row << bwl);
    float b;
    if (c == eob - 1) {
        const int bctx = c == 0 ? 0 : c <= (height << bwl) / 8 ? 1 : c <= (height << bwl) / 4 ? 2 : 3;
        b = C.coeff_base_eob[ctx][ptype][bctx][std::min(level, 3) - 1];
    } else {
        int mag = 0;
        for (int k = 0; k < 5; k++) {
            const int rr = row + av1t::Sig_Ref_Diff_Offset[txClass][k][0], cc = col + av1t::Sig_Ref_Diff_Offset[txClass][k][1];
            if (rr >= 0 && cc >= 0 && rr < height && cc < width) { const int a = q[(rr << bwl) + cc]; mag += std::min(a < 0 ? -a : a, 3); }
        }
        int bctx = std::min((mag + 1) >> 1, 4);
        if (txClass == TX_CLASS_2D) {
            if (row == 0 && col == 0) bctx = 0;
            else bctx += av1t::Coeff_Base_Ctx_Offset[txSz][std::min(row, 4)][std::min(col, 4)];
        } else {
            const int id = txClass == TX_CLASS_VERT ? row : col;
            bctx += av1t::Coeff_Base_Pos_Ctx_Offset[std::min(id, 2)];
        }
        b = C.coeff_base[ctx][ptype][bctx][std::min(level, 3)];
    }
    if (level > 2) {
        int mag = 0;
        for (int k = 0; k < 3; k++) {
            const int rr = row + av1t::Mag_Ref_Offset_With_Tx_Class[txClass][k][0], cc = col + av1t::Mag_Ref_Offset_With_Tx_Class[txClass][k][1];
            if (rr >= 0 && cc >= 0 && rr < height && cc < width) { const int a = q[rr * width + cc]; mag += std::min(a < 0 ? -a : a, 15); }
        }
        mag = std::min((mag + 1) >> 1, 6);
        int brctx;
        if (pos == 0) brctx = mag;
        else if (txClass == TX_CLASS_2D) brctx = (row < 2 && col < 2) ? mag + 7 : mag + 14;
        else if (txClass == TX_CLASS_HORIZ) brctx = col == 0 ? mag + 7 : mag + 14;
        else brctx = row == 0 ? mag + 7 : mag + 14;
        const int brsz = std::min(ctx, (int)TX_32X32);
        int rem = level - 3;
        for (int i = 0; i < 4; i++) {
            const int k = std::min(rem, 3);
            b += C.coeff_br[brsz][ptype][brctx][k];
            rem -= k;
            if (k < 3) break;
        }
    }
    if (v) {
        if (c == 0) b += C.dc_sign[ptype][dcCtx][qv < 0];
        else b += 1.0f;
        if (v > 14) {
            int len = 0, xg = v - 14;
            while (xg > 1) { xg >>= 1; len++; }
            b += (float)(2 * len + 1);
        }
    }
    return b;
}

float res_energy(const int16_t *res, int n) {   // k_analysis.cuh:326-333 (raster order, 32 virtual lanes)
    float e[32] = {0};
    for (int p = 0; p < n; p++) e[p & 31] += (float)(res[p] * res[p]);
    return lane_sum(e);
}

float rate_of_levels(const int16_t *q, const EvalIn &e, int *eobOut) {   // dev_tx.cuh:301-316
    const int t = e.txSz, n = std::min(txw(t), 32) * std::min(txh(t), 32);
    const uint16_t *scan = get_scan(t, e.txType);
    int eob = 0;
    for (int c = 0; c < n; c++) if (q[scan[c]]) eob = c + 1;
    float bits[32] = {0};
    bits[0] = base_bits(e, eob);
    const int cls = tx_class(e.txType);
    for (int c = 0; c < eob; c++) bits[c & 31] += coeff_bits(c, eob, q, scan, t, cls, e.plane > 0, e.dcCtx, *e.C);
    *eobOut = eob;
    return lane_sum(bits);
}

// Work accounting, by its definition: the forward-transform FLOPs of one evaluation are 2 per multiply-add of the
// column pass (W columns x min(H, 32) kept outputs x H taps) and of the row pass (min(H, 32) rows x min(W, 32) kept
// outputs x W taps), 1 per multiply where a pass is an identity kind, and 1 per coefficient for the row scale.
static uint64_t fwd_flops(const EvalIn &e) {
    const uint64_t W = txw(e.txSz), H = txh(e.txSz), tw = W > 32 ? 32 : W, th = H > 32 ? 32 : H;
    const uint64_t col = col_kind(e.txType) == K_IDT ? W * th : 2 * W * th * H;
    const uint64_t row = row_kind(e.txType) == K_IDT ? th * tw : 2 * th * tw * W;
    return col + row + th * tw;
}
static void count_eval(Frame &F, int phase, const EvalIn &e) { F.flops[phase].fetch_add(fwd_flops(e), std::memory_order_relaxed); }

EvalOut eval_txb(const int16_t *res, const EvalIn &e, int16_t *q, float *coefOut) {
    const Tables &T = tables();
    const int t = e.txSz, W = txw(t), H = txh(t), tw = W > 32 ? 32 : W, th = H > 32 ? 32 : H;
    const int lw = av1t::Tx_Width_Log2[t], lh = av1t::Tx_Height_Log2[t];
    const int rk = row_kind(e.txType), ck = col_kind(e.txType), fud = flip_ud(e.txType), flr = flip_lr(e.txType);
    const float *Fc = T.fwd[ck][lh].data(), *Fr = T.fwd[rk][lw].data();
    static thread_local float tmp[32 * 64], coef[32 * 32];
    // column pass (dev_tx.cuh:171-207)
    for (int j = 0; j < W; j++) {
        const int jj = flr ? W - 1 - j : j;
        for (int i = 0; i < th; i++) {
            float acc;
            if (ck == K_IDT) acc = Fc[i * H + i] * (float)res[(fud ? H - 1 - i : i) * W + jj];
            else {
                acc = 0.f;
                for (int k = 0; k < H; k++) acc = fmaf(Fc[i * H + k], (float)res[(fud ? H - 1 - k : k) * W + jj], acc);
            }
            tmp[i * W + j] = acc;
        }
    }
    // row pass (dev_tx.cuh:208-238)
    const float scale = T.txScale[t];
    for (int i = 0; i < th; i++)
        for (int jj = 0; jj < tw; jj++) {
            float acc;
            if (rk == K_IDT) acc = Fr[jj * W + jj] * tmp[i * W + jj];
            else {
                acc = 0.f;
                for (int k = 0; k < W; k++) acc = fmaf(Fr[jj * W + k], tmp[i * W + k], acc);
            }
            coef[i * tw + jj] = acc * scale;
        }
    // quantization + distortion (dev_tx.cuh:271-300)
    const int n = tw * th, den = dq_denom(t);
    float dist[32] = {0}, kept[32] = {0};
    for (int idx = 0; idx < n; idx++) {
        const float v = coef[idx];
        const int step = idx == 0 ? e.dcQ : e.acQ;
        const float a = fabsf(v) * (float)den / (float)step;
        int l = a < e.zbin ? 0 : (int)(a + e.rnd);
        l = std::min(l, 16383);
        const int qv = v < 0 ? -l : l;
        q[idx] = (int16_t)qv;
        const float dq = l ? (float)dequant_coef(qv, step, den, e.bd) : 0.f;
        const float d = v - dq;
        const float sq = d * d;
        dist[idx & 31] = dist[idx & 31] + sq;
        kept[idx & 31] = fmaf(v, v, kept[idx & 31]);
    }
    if (coefOut) memcpy(coefOut, coef, sizeof(float) * n);
    const float distS = lane_sum(dist);
    EvalOut o;
    o.R = rate_of_levels(q, e, &o.eob);
    const float g = T.txg[t][e.txType];
    o.D = g * distS;
    if (W == 64 || H == 64) {
        const float keptS = lane_sum(kept);
        o.D = o.D + fmaxf(0.f, fmaf(-g, keptS, e.resEnergy));
    }
    return o;
}

// ------------------------------------------------------------------------------------------------------------
// frame setup (pipeline.cpp:329-420, 433-451)
// ------------------------------------------------------------------------------------------------------------
static int tile_log2(int blk, int target) { int k = 0; while ((blk << k) < target) k++; return k; }
int tile_layout(int W, int H, int *cols, int *rows, int *colStart, int *rowStart) {
    const int MiCols = 2 * ((W + 7) >> 3), MiRows = 2 * ((H + 7) >> 3);
    const int sbCols = (MiCols + 15) >> 4, sbRows = (MiRows + 15) >> 4;
    const int maxW = 4096 >> 6, maxA = (4096 * 2304) >> 12;
    const int minC = tile_log2(maxW, sbCols);
    const int maxR = tile_log2(1, std::min(sbRows, 64));
    int minT = tile_log2(maxA, sbRows * sbCols);
    if (minT < minC) minT = minC;
    const int cl = minC, tws = (sbCols + (1 << cl) - 1) >> cl;
    int i = 0;
    for (int s = 0; s < sbCols; s += tws) colStart[i++] = s << 4;
    colStart[i] = MiCols;
    *cols = i;
    int rl = std::max(minT - cl, 0);
    for (;;) {
        if (rl > maxR) return -1;
        const int ths = (sbRows + (1 << rl) - 1) >> rl;
        if (tws * ths <= maxA) {
            i = 0;
            for (int s = 0; s < sbRows; s += ths) rowStart[i++] = s << 4;
            rowStart[i] = MiRows;
            *rows = i;
            return 0;
        }
        rl++;
    }
}

void frame_setup(Frame &F, int W, int H, int ssx, int mono, int bd, int qindex, int edgeFilter, const uint16_t *const src[3],
                 const int srcStride[3]) {
    F.W = W; F.H = H;
    F.MiCols = 2 * ((W + 7) >> 3); F.MiRows = 2 * ((H + 7) >> 3);
    F.sbCols = (F.MiCols + 15) >> 4; F.sbRows = (F.MiRows + 15) >> 4;
    F.aw = F.sbCols * 64; F.ah = F.sbRows * 64;
    F.ssx = F.ssy = mono ? 1 : ssx;
    F.mono = mono; F.bd = bd; F.qindex = qindex; F.edgeFilter = edgeFilter;
    F.qctx = qindex <= 20 ? 0 : qindex <= 60 ? 1 : qindex <= 120 ? 2 : 3;   // cdfctx.h:88-93
    tile_layout(W, H, &F.tileCols, &F.tileRows, F.tileColStart, F.tileRowStart);
    // source planes padded by edge replication to the 64-aligned area (the caller's planes hold at least W x H)
    for (int p = 0; p < 3; p++) {
        const int sx = p ? F.ssx : 0, sy = p ? F.ssy : 0;
        const int pw = F.aw >> sx, ph = F.ah >> sy, vw = (W + sx) >> sx, vh = (H + sy) >> sy;
        F.stride[p] = pw;
        F.src[p].assign((size_t)pw * ph, (uint16_t)(1 << (bd - 1)));
        F.rec[p].assign((size_t)pw * ph, 0);
        if (p && mono) continue;
        for (int y = 0; y < ph; y++)
            for (int x = 0; x < pw; x++) F.src[p][(size_t)y * pw + x] = src[p][(size_t)std::min(y, vh - 1) * srcStride[p] + std::min(x, vw - 1)];
    }
    const size_t mi = (size_t)F.MiCols * F.MiRows;
    F.mY.assign(mi, 0); F.mUV.assign(mi, 0); F.mSkip.assign(mi, 0); F.mMiSize.assign(mi, 0); F.mTx.assign(mi, 0);
    for (int p = 0; p < 3; p++) {
        const size_t n = p ? (size_t)(F.MiCols >> F.ssx) * (F.MiRows >> F.ssy) : mi;
        F.cLvl[p].assign(n, 0);
        F.cDc[p].assign(n, 0);
    }
    for (int s = 0; s < 22; s++) {
        F.nbx[s] = F.nby[s] = 0;
        F.blk[s].clear();
        if (s >= 13 && s <= 15) continue;
        const int w = 4 * bw4(s), h = 4 * bh4(s);
        F.nbx[s] = (F.MiCols * 4 + w - 1) / w;
        F.nby[s] = (F.MiRows * 4 + h - 1) / h;
        F.blk[s].assign((size_t)F.nbx[s] * F.nby[s], BlockRes{INF, {0}, 0, 0});
    }
    const int nsb = F.sbCols * F.sbRows;
    F.leaves.assign(nsb, {});
    F.sbJ.assign(nsb, 0.f);
    F.sbBlocks.assign(nsb, {}); F.sbBlockTxb.assign(nsb, {}); F.sbTxb.assign(nsb, {}); F.sbCoef.assign(nsb, {});
    memset(F.stats, 0, sizeof F.stats);
}

struct TileBox { int x0, y0, x1, y1; };
static TileBox tile_of(const Frame &F, int x, int y) {   // dev_common.cuh:133-143
    int c = 0, r = 0;
    while (c + 1 < F.tileCols && x >= F.tileColStart[c + 1] * 4) c++;
    while (r + 1 < F.tileRows && y >= F.tileRowStart[r + 1] * 4) r++;
    return TileBox{F.tileColStart[c] * 4, F.tileRowStart[r] * 4, F.tileColStart[c + 1] * 4, F.tileRowStart[r + 1] * 4};
}
static int zorder16(int x, int y) {
    int z = 0;
    for (int b = 0; b < 4; b++) z |= (((x >> b) & 1) << (2 * b)) | (((y >> b) & 1) << (2 * b + 1));
    return z;
}
// open-loop guess of "(px, py) is decoded before the block at (x, y)" (dev_common.cuh:151-163), luma
static bool ol_decoded(int px, int py, int x, int y, const TileBox &t) {
    if (px < t.x0 || py < t.y0 || px >= t.x1 || py >= t.y1) return false;
    const int sbx = (x / 64) * 64, sby = (y / 64) * 64;
    if (py < sby) return true;
    if (px < sbx) return py < sby + 64;
    if (px >= sbx + 64 || py >= sby + 64) return false;
    return zorder16(((px - sbx) / 4) & 15, ((py - sby) / 4) & 15) < zorder16(((x - sbx) / 4) & 15, ((y - sby) / 4) & 15);
}

static inline int cand_mode(int c) { return c >> 3; }
static inline int cand_delta(int c) { return (c & 7) - 3; }
static inline int cand_pack(int m, int d) { return (m << 3) | (d + 3); }
static int tx_of_shape(int bs) { return av1t::Max_Tx_Size_Rect[bs]; }   // one luma tx block of the block's size

// sum |4x4 Hadamard coefficients| of d (k_analysis.cuh:47-63)
static int hadamard4_sum(int *d) {
    for (int i = 0; i < 4; i++) {
        int *r = d + 4 * i;
        const int s0 = r[0] + r[1], s1 = r[0] - r[1], s2 = r[2] + r[3], s3 = r[2] - r[3];
        r[0] = s0 + s2; r[1] = s1 + s3; r[2] = s0 - s2; r[3] = s1 - s3;
    }
    int sum = 0;
    for (int j = 0; j < 4; j++) {
        const int s0 = d[j] + d[4 + j], s1 = d[j] - d[4 + j], s2 = d[8 + j] + d[12 + j], s3 = d[8 + j] - d[12 + j];
        sum += abs(s0 + s2) + abs(s1 + s3) + abs(s0 - s2) + abs(s1 - s3);
    }
    return sum;
}
// SATD (integer total of all 4x4 sub-blocks) of prediction pred against src (w x h)
static int satd_sum(const uint16_t *src, int sp, const int *pred, int w, int h) {
    int t = 0;
    for (int sy = 0; sy < h; sy += 4)
        for (int sx = 0; sx < w; sx += 4) {
            int d[16];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) d[i * 4 + j] = (int)src[(sy + i) * sp + sx + j] - pred[(sy + i) * w + sx + j];
            t += hadamard4_sum(d);
        }
    return t;
}

// ------------------------------------------------------------------------------------------------------------
// open-loop analysis of one shape (k_analysis.cuh:496-689; k_analysis_small.cuh:123-311)
// ------------------------------------------------------------------------------------------------------------
static void analyse_block(Frame &F, const Params &P, int bs, int bx, int by) {
    const Tables &T = tables();
    const int W = 4 * bw4(bs), H = 4 * bh4(bs), lw = av1t::Mi_Width_Log2[bs] + 2, lh = av1t::Mi_Height_Log2[bs] + 2;
    const int x = bx * W, y = by * H;
    const size_t bi = (size_t)by * F.nbx[bs] + bx;
    BlockRes &out = F.blk[bs][bi];
    if (x >= F.MiCols * 4 || y >= F.MiRows * 4) { out.J = INF; out.ncand = 0; return; }
    if (F.analyse[bs] && !F.analyse[bs][bi]) { out.J = INF; out.ncand = 0; return; }   // not in the shape's block list
    const int txSz = tx_of_shape(bs), set = tx_set_intra(txSz), bd = F.bd;
    const TileBox tb = tile_of(F, x, y);
    const int haveL = x > tb.x0, haveA = y > tb.y0;
    const int haveAR = ol_decoded(x + W, y - 1, x, y, tb), haveBL = ol_decoded(x - 1, y + H, x, y, tb);
    const int maxX = F.MiCols * 4 - 1, maxY = F.MiRows * 4 - 1;
    const int sp = F.stride[0];
    const uint16_t *src = F.src[0].data() + (size_t)y * sp + x;
    const float lambda = T.lambda[bd == 10][F.qindex] * P.lambdaMul, lamS = P.satdLambda * T.lamS[bd == 10][F.qindex];
    const CostTab &C = T.cost[F.qctx];
    uint16_t ea[EDGE_LEN], el[EDGE_LEN];
    uint16_t *A = ea + EDGE_OFF, *L = el + EDGE_OFF;
    build_edges(F.src[0].data(), sp, x, y, lw, lh, haveL, haveA, haveAR, haveBL, maxX, maxY, A, L, bd);
    const int dcv = dc_value(A, L, haveL, haveA, lw, lh, bd);
    auto mbits = [&](int c) -> float {
        const int m = cand_mode(c);
        float b = T.avgYMode[m];
        if (bs >= BLOCK_8X8 && is_directional(m)) b += C.angle_delta[m - V_PRED][cand_delta(c) + 3];
        return b;
    };
    static thread_local int pred[64 * 64];
    // screening: unfiltered edges for every candidate (k_analysis.cuh:189-221)
    auto screen = [&](int c) -> float {
        predict_block(cand_mode(c), cand_delta(c), 0, 0, haveL, haveA, lw, lh, x, y, maxX, maxY, A, L, dcv, pred, bd);
        return (float)satd_sum(src, sp, pred, W, H) * 0.25f;
    };
    int cands[32];
    float js[32];
    int nc = 13;
    const int K = std::min(P.K, MAXK);
    int top[MAXK], nK = 0;
    // an external screener's candidates, best first (k_analysis.cuh:586-600): entries with a mode of 13 or more are
    // skipped, deltas kept for directional modes of blocks from 8x8 (outside -3..3: 0); a block with no usable entry is
    // screened as without the hook
    if (F.ext[bs]) {
        const ExtCand *ec = F.ext[bs] + bi * F.extTopK;
        for (int k = 0; k < F.extTopK && nK < MAXK; k++) {
            const int m = ec[k].mode;
            if (m >= 13) continue;
            const int d = (bs >= BLOCK_8X8 && is_directional(m) && ec[k].delta >= -3 && ec[k].delta <= 3) ? ec[k].delta : 0;
            cands[nK] = cand_pack(m, d);
            top[nK] = nK;
            nK++;
        }
        nK = std::min(nK, K);
    }
    if (!nK) {
    for (int m = 0; m < 13; m++) { cands[m] = cand_pack(m, 0); js[m] = fmaf(lamS, mbits(cands[m]), screen(cands[m])); }
    if (P.deltas && bs >= BLOCK_8X8) {
        int b0 = -1, b1 = -1;   // the two best directional modes, first minimum on ties
        for (int m = V_PRED; m <= D67_PRED; m++) if (b0 < 0 || js[m] < js[b0]) b0 = m;
        for (int m = V_PRED; m <= D67_PRED; m++) if (m != b0 && (b1 < 0 || js[m] < js[b1])) b1 = m;
        for (int k = 0; k < 12; k++) {
            const int m = k < 6 ? b0 : b1, kk = k % 6, d = kk < 3 ? -(kk + 1) : kk - 2;
            cands[13 + k] = cand_pack(m, d);
            js[13 + k] = fmaf(lamS, mbits(cands[13 + k]), screen(cands[13 + k]));
        }
        nc = 25;
    }
    bool used[32] = {false};
    for (int k = 0; k < K; k++) {
        int b = -1;
        for (int i = 0; i < nc; i++) if (!used[i] && js[i] < INF && (b < 0 || js[i] < js[b])) b = i;
        if (b < 0) break;
        used[b] = true;
        top[nK++] = b;
    }
    }
    // RD: top-K with DCT_DCT, then the other tx types of the set for the winner (k_analysis.cuh:641-677)
    const int hasC = has_chroma(bs, y >> 2, x >> 2, F.mono, F.ssx, F.ssy);
    const float overhead = (hasC ? P.uvOverheadBits : 0.f) + T.avgSkip0;
    EvalIn e;
    e.txSz = txSz; e.plane = 0; e.dcQ = dc_q(F.qindex, bd); e.acQ = ac_q(F.qindex, bd); e.rnd = P.rnd; e.zbin = P.zbin;
    e.skipCtx = 0; e.dcCtx = 0; e.bd = bd; e.C = &C; e.resEnergy = 0.f;
    static thread_local int16_t res[64 * 64], q[32 * 32];
    float bestJ = INF;
    int bestK = 0, bestTx = DCT_DCT, resK = -1;
    const int nSteps = nK ? nK + ((P.txSearch && set != TX_SET_DCTONLY) ? 15 : 0) : 0;
    for (int st = 0; st < nSteps; st++) {
        const int k = st < nK ? st : bestK, t = st < nK ? DCT_DCT : st - nK + 1;
        if (t != DCT_DCT && !tx_in_set(set, t)) continue;
        const int c = cands[top[k]];
        if (k != resK) {
            predict_block(cand_mode(c), cand_delta(c), F.edgeFilter, 0, haveL, haveA, lw, lh, x, y, maxX, maxY, A, L, dcv, pred, bd);
            for (int i = 0; i < H; i++)
                for (int j = 0; j < W; j++) res[i * W + j] = (int16_t)((int)src[i * sp + j] - pred[i * W + j]);
            e.resEnergy = (W == 64 || H == 64) ? res_energy(res, W * H) : 0.f;
            resK = k;
        }
        e.txType = t;
        e.intraDir = cand_mode(c);
        const EvalOut o = eval_txb(res, e, q);
        count_eval(F, 0, e);
        const float J = fmaf(lambda, (o.R + mbits(c)) + overhead, o.D);
        if (J < bestJ) { bestJ = J; if (st < nK) bestK = k; else bestTx = t; }
    }
    out.J = nK ? bestJ : INF;
    out.tx = (uint8_t)bestTx;
    int n = 0;
    if (nK) out.cand[n++] = (uint8_t)cands[top[bestK]];
    for (int k = 0; k < nK; k++) if (k != bestK) out.cand[n++] = (uint8_t)cands[top[k]];
    out.ncand = (uint8_t)n;
}

void analysis(Frame &F, const Params &P, int bs) {
    if (bs >= BLOCK_INVALID || (bs >= 13 && bs <= 15) || !((P.shapeMask >> bs) & 1)) return;
    for (int by = 0; by < F.nby[bs]; by++)
        for (int bx = 0; bx < F.nbx[bs]; bx++) analyse_block(F, P, bs, bx, by);
}

// ------------------------------------------------------------------------------------------------------------
// partition DP (k_dp.cuh:72-297); the shapes of the plain search always leave a partition of finite J
// ------------------------------------------------------------------------------------------------------------
static int sq_bsize(int n) { return n == 4 ? BLOCK_4X4 : n == 8 ? BLOCK_8X8 : n == 16 ? BLOCK_16X16 : n == 32 ? BLOCK_32X32 : BLOCK_64X64; }
static int sub_bsize(int p, int n) { return av1t::Partition_Subsize[p][sq_bsize(n)]; }

static void dp_sb(Frame &F, const Params &P, int sb) {
    const Tables &T = tables();
    const int sbx = (sb % F.sbCols) * 64, sby = (sb / F.sbCols) * 64;
    const int FW = F.MiCols * 4, FH = F.MiRows * 4;
    const float lambda = T.lambda[F.bd == 10][F.qindex] * P.lambdaMul;
    auto blockJ = [&](int bs, int x, int y) -> float {
        if (bs >= BLOCK_INVALID || F.nbx[bs] == 0 || !((P.shapeMask >> bs) & 1)) return INF;
        const int w = 4 * bw4(bs), h = 4 * bh4(bs);
        return F.blk[bs][(size_t)(y / h) * F.nbx[bs] + x / w].J;
    };
    struct Node { float J; int p; } node[85];
    const int lvBase[4] = {0, 1, 5, 21};
    for (int L = 3; L >= 0; L--) {
        const int n = 64 >> L, half = n >> 1, q4 = n >> 2, per = 1 << L;
        const int bsl = L == 3 ? 1 : L == 2 ? 2 : L == 1 ? 3 : 4;
        const int pmask = n == 8 ? P.partMask8 : P.partMaskN;
        for (int ny = 0; ny < per; ny++)
            for (int nx = 0; nx < per; nx++) {
                Node &nd = node[lvBase[L] + ny * per + nx];
                const int x = sbx + nx * n, y = sby + ny * n;
                if (x >= FW || y >= FH) { nd.J = 0.f; nd.p = -1; continue; }
                const bool hasRows = y + half < FH, hasCols = x + half < FW;
                float splitJ = 0.f;
                if (n == 8) {
                    for (int k = 0; k < 4; k++) {
                        const int xx = x + (k & 1) * 4, yy = y + (k >> 1) * 4;
                        if (xx < FW && yy < FH) splitJ = splitJ + blockJ(BLOCK_4X4, xx, yy);
                    }
                } else {
                    const int cb = lvBase[L + 1], cper = per * 2;
                    for (int k = 0; k < 4; k++) splitJ = splitJ + node[cb + (ny * 2 + (k >> 1)) * cper + nx * 2 + (k & 1)].J;
                }
                float best = INF;
                int bp = PARTITION_SPLIT;
                const float lamN = lambda * 1.0f;
                auto consider = [&](int p, float J, float bits) {
                    const float t = fmaf(lamN, bits, J);
                    if (t < best) { best = t; bp = p; }
                };
                const bool big = n > P.maxPart;
                if (hasRows && hasCols) {
                    const float *pb = T.avgPart[bsl];
                    consider(PARTITION_SPLIT, splitJ, pb[PARTITION_SPLIT]);
                    if (!big) {
                        const int hz = sub_bsize(PARTITION_HORZ, n), vt = sub_bsize(PARTITION_VERT, n), sp = sub_bsize(PARTITION_SPLIT, n);
                        if (pmask & (1 << PARTITION_NONE)) {
                            const float jn = blockJ(sq_bsize(n), x, y);
                            if (jn < INF) consider(PARTITION_NONE, jn * P.splitBias, pb[PARTITION_NONE]);
                        }
                        if (pmask & (1 << PARTITION_HORZ)) consider(PARTITION_HORZ, blockJ(hz, x, y) + blockJ(hz, x, y + half), pb[PARTITION_HORZ]);
                        if (pmask & (1 << PARTITION_VERT)) consider(PARTITION_VERT, blockJ(vt, x, y) + blockJ(vt, x + half, y), pb[PARTITION_VERT]);
                        if (n > 8) {
                            const float s00 = blockJ(sp, x, y), s01 = blockJ(sp, x + half, y), s10 = blockJ(sp, x, y + half),
                                        s11 = blockJ(sp, x + half, y + half);
                            if (pmask & (1 << PARTITION_HORZ_A)) consider(PARTITION_HORZ_A, s00 + s01 + blockJ(hz, x, y + half), pb[PARTITION_HORZ_A]);
                            if (pmask & (1 << PARTITION_HORZ_B)) consider(PARTITION_HORZ_B, blockJ(hz, x, y) + s10 + s11, pb[PARTITION_HORZ_B]);
                            if (pmask & (1 << PARTITION_VERT_A)) consider(PARTITION_VERT_A, s00 + s10 + blockJ(vt, x + half, y), pb[PARTITION_VERT_A]);
                            if (pmask & (1 << PARTITION_VERT_B)) consider(PARTITION_VERT_B, blockJ(vt, x, y) + s01 + s11, pb[PARTITION_VERT_B]);
                            const int h4 = sub_bsize(PARTITION_HORZ_4, n), v4 = sub_bsize(PARTITION_VERT_4, n);
                            if (pmask & (1 << PARTITION_HORZ_4)) {
                                float J = 0.f;
                                for (int i = 0; i < 4; i++) if (i < 3 || y + 3 * q4 < FH) J = J + blockJ(h4, x, y + i * q4);
                                consider(PARTITION_HORZ_4, J, pb[PARTITION_HORZ_4]);
                            }
                            if (pmask & (1 << PARTITION_VERT_4)) {
                                float J = 0.f;
                                for (int i = 0; i < 4; i++) if (i < 3 || x + 3 * q4 < FW) J = J + blockJ(v4, x + i * q4, y);
                                consider(PARTITION_VERT_4, J, pb[PARTITION_VERT_4]);
                            }
                        }
                    }
                } else if (hasCols) {
                    consider(PARTITION_SPLIT, splitJ, 1.f);
                    if (!big) consider(PARTITION_HORZ, blockJ(sub_bsize(PARTITION_HORZ, n), x, y), 1.f);
                } else if (hasRows) {
                    consider(PARTITION_SPLIT, splitJ, 1.f);
                    if (!big) consider(PARTITION_VERT, blockJ(sub_bsize(PARTITION_VERT, n), x, y), 1.f);
                } else {
                    consider(PARTITION_SPLIT, splitJ, 0.f);
                }
                nd.J = best;
                nd.p = bp;
            }
    }
    F.sbJ[sb] = node[0].J;
    std::vector<Leaf> &out = F.leaves[sb];
    out.clear();
    auto leaf = [&](int bs, int x, int y) {
        if (x >= FW || y >= FH) return;
        out.push_back(Leaf{(uint8_t)bs, (uint8_t)((x - sbx) >> 2), (uint8_t)((y - sby) >> 2), 0});
    };
    // pre-order walk (decisions.h:71-91)
    struct Item { int L, nx, ny; };
    std::vector<Item> stk;
    stk.push_back({0, 0, 0});
    while (!stk.empty()) {
        const Item it = stk.back();
        stk.pop_back();
        const int n = 64 >> it.L, half = n >> 1, q4 = n >> 2, per = 1 << it.L;
        const Node &nd = node[lvBase[it.L] + it.ny * per + it.nx];
        const int x = sbx + it.nx * n, y = sby + it.ny * n;
        if (nd.p < 0) continue;
        const int p = nd.p, sub = sub_bsize(p, n), sp2 = sub_bsize(PARTITION_SPLIT, n);
        const bool hasRows = y + half < FH, hasCols = x + half < FW;
        switch (p) {
        case PARTITION_NONE: leaf(sub, x, y); break;
        case PARTITION_HORZ: leaf(sub, x, y); if (hasRows) leaf(sub, x, y + half); break;
        case PARTITION_VERT: leaf(sub, x, y); if (hasCols) leaf(sub, x + half, y); break;
        case PARTITION_SPLIT:
            if (n == 8) { leaf(BLOCK_4X4, x, y); leaf(BLOCK_4X4, x + 4, y); leaf(BLOCK_4X4, x, y + 4); leaf(BLOCK_4X4, x + 4, y + 4); }
            else for (int k = 3; k >= 0; k--) stk.push_back({it.L + 1, it.nx * 2 + (k & 1), it.ny * 2 + (k >> 1)});
            break;
        case PARTITION_HORZ_A: leaf(sp2, x, y); leaf(sp2, x + half, y); leaf(sub, x, y + half); break;
        case PARTITION_HORZ_B: leaf(sub, x, y); leaf(sp2, x, y + half); leaf(sp2, x + half, y + half); break;
        case PARTITION_VERT_A: leaf(sp2, x, y); leaf(sp2, x, y + half); leaf(sub, x + half, y); break;
        case PARTITION_VERT_B: leaf(sub, x, y); leaf(sp2, x + half, y); leaf(sp2, x + half, y + half); break;
        case PARTITION_HORZ_4: for (int i = 0; i < 4; i++) if (i < 3 || y + 3 * q4 < FH) leaf(sub, x, y + i * q4); break;
        default: for (int i = 0; i < 4; i++) if (i < 3 || x + 3 * q4 < FW) leaf(sub, x + i * q4, y); break;
        }
    }
}
void partition_dp(Frame &F, const Params &P) {
    for (int sb = 0; sb < F.sbCols * F.sbRows; sb++) dp_sb(F, P, sb);
}

// ------------------------------------------------------------------------------------------------------------
// closed loop of one SB (k_recon.cuh)
// ------------------------------------------------------------------------------------------------------------
struct SbState {
    uint8_t bdm[3][18][18];   // BlockDecoded[plane][y + 1][x + 1], y, x in -1 .. 16
    int nTxb, nCoef;
};

// coefficient contexts of a tx block from the 4x4 context maps (k_recon.cuh:203-238; spec get_tx_size_context / dc ctx)
static void txb_ctx(const Frame &F, int plane, int txSz, int x4, int y4, int bsize, const TileBox &tb, int *skipCtx, int *dcCtx) {
    const int sx = plane ? F.ssx : 0, sy = plane ? F.ssy : 0;
    const int stride = F.MiCols >> sx, maxX4 = F.MiCols >> sx, maxY4 = F.MiRows >> sy;
    const uint8_t *L = F.cLvl[plane].data(), *D = F.cDc[plane].data();
    const bool hasA = y4 > ((tb.y0 >> 2) >> sy), hasL = x4 > ((tb.x0 >> 2) >> sx);
    const int w = txw(txSz), h = txh(txSz), w4 = w >> 2, h4 = h >> 2;
    int top = 0, left = 0, topAny = 0, leftAny = 0, dcSign = 0;
    if (hasA)
        for (int k = 0; k < w4 && x4 + k < maxX4; k++) {
            const int l = L[(size_t)(y4 - 1) * stride + x4 + k], d = D[(size_t)(y4 - 1) * stride + x4 + k];
            top = std::max(top, l);
            topAny |= l | d;
            dcSign += d == 2 ? 1 : d == 1 ? -1 : 0;
        }
    if (hasL)
        for (int k = 0; k < h4 && y4 + k < maxY4; k++) {
            const int l = L[(size_t)(y4 + k) * stride + x4 - 1], d = D[(size_t)(y4 + k) * stride + x4 - 1];
            left = std::max(left, l);
            leftAny |= l | d;
            dcSign += d == 2 ? 1 : d == 1 ? -1 : 0;
        }
    if (plane == 0) {
        const int bw = 4 * bw4(bsize), bh = 4 * bh4(bsize);
        if (bw == w && bh == h) *skipCtx = 0;
        else if (top == 0 && left == 0) *skipCtx = 1;
        else if (top == 0 || left == 0) *skipCtx = 2 + (std::max(top, left) > 3);
        else if (std::max(top, left) <= 3) *skipCtx = 4;
        else if (std::min(top, left) <= 3) *skipCtx = 5;
        else *skipCtx = 6;
    } else {
        const int ub = uv_bsize(bsize, F.ssx, F.ssy);
        *skipCtx = (topAny != 0) + (leftAny != 0) + 7 + ((4 * bw4(ub) * 4 * bh4(ub) > w * h) ? 3 : 0);
    }
    *dcCtx = dcSign < 0 ? 1 : dcSign > 0 ? 2 : 0;
}

// exact dequant + inverse transform + add of one tx block into rec, with the reference's range repair (halving the
// levels until the inverse transform is conformant, k_recon.cuh:139-197).  Returns the number of halvings.
static int inverse_add(const Frame &F, int16_t *q, int txSz, int txType, int dcQ, int acQ, uint16_t *dst, int pitch) {
    const int w = txw(txSz), h = txh(txSz), n = tx_coeffs(txSz), den = dq_denom(txSz);
    static thread_local int32_t dq[1024], res[4096];
    int fixes = 0;
    for (;;) {
        for (int i = 0; i < n; i++) dq[i] = q[i] ? dequant_coef(q[i], i == 0 ? dcQ : acQ, den, F.bd) : 0;
        uint32_t err = 0;
        inv_txfm2d(dq, txSz, txType, res, &err, F.bd);
        if (!err) break;
        fixes++;
        bool nz = false;
        for (int i = 0; i < n; i++) { q[i] = (int16_t)(q[i] / 2); nz |= q[i] != 0; }
        if (!nz) return fixes;
    }
    const int fud = flip_ud(txType), flr = flip_lr(txType);
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            uint16_t *d = dst + (fud ? h - 1 - i : i) * pitch + (flr ? w - 1 - j : j);
            *d = (uint16_t)clip1((int)*d + res[i * w + j], F.bd);
        }
    return fixes;
}

// record a finished tx block (k_recon.cuh:685-730): coefficients, txb entry, context maps, BlockDecoded
static int record_txb(Frame &F, SbState &S, int sb, int plane, int txSz, int txType, int x4, int y4, const int16_t *q) {
    const int n = tx_coeffs(txSz);
    const uint16_t *scan = get_scan(txSz, txType);
    int eob = 0;
    for (int c = 0; c < n; c++) if (q[scan[c]]) eob = c + 1;
    int cul = 0;
    for (int c = 0; c < eob; c++) cul += abs(q[scan[c]]);
    cul = std::min(cul, 63);
    const int dcCat = q[0] < 0 ? 1 : q[0] > 0 ? 2 : 0;
    std::vector<int16_t> &co = F.sbCoef[sb];
    const uint32_t off = (uint32_t)co.size();
    co.insert(co.end(), q, q + n);
    const int sx = plane ? F.ssx : 0, sy = plane ? F.ssy : 0;
    const int stride = F.MiCols >> sx, rows = F.MiRows >> sy, w4 = txw(txSz) >> 2, h4 = txh(txSz) >> 2;
    for (int i = 0; i < h4; i++)
        for (int j = 0; j < w4; j++)
            if (y4 + i < rows && x4 + j < stride) {
                F.cLvl[plane][(size_t)(y4 + i) * stride + x4 + j] = (uint8_t)(eob ? cul : 0);
                F.cDc[plane][(size_t)(y4 + i) * stride + x4 + j] = (uint8_t)(eob ? dcCat : 0);
            }
    const int r4 = y4 & ((16 >> sy) - 1), c4 = x4 & ((16 >> sx) - 1);   // SB-relative plane 4x4 units
    for (int i = 0; i < h4; i++)
        for (int j = 0; j < w4; j++)
            if (r4 + i <= 16 && c4 + j <= 16) S.bdm[plane][r4 + i + 1][c4 + j + 1] = 1;
    RecTxb t;
    t.coeff_off = off; t.x4 = (uint16_t)x4; t.y4 = (uint16_t)y4; t.plane = (uint8_t)plane; t.tx_size = (uint8_t)txSz;
    t.tx_type = (uint8_t)(plane == 0 && eob == 0 ? DCT_DCT : txType);
    t.flags = 0; t.eob = (uint16_t)eob; t.reserved = 0;
    F.sbTxb[sb].push_back(t);
    S.nTxb++;
    S.nCoef += n;
    return eob;
}

struct PlaneEdges {
    uint16_t ea[EDGE_LEN], el[EDGE_LEN];
    int dcv, haveL, haveA, x, y, maxX, maxY, ft, lw, lh;
    const uint16_t *A() const { return ea + EDGE_OFF; }
    const uint16_t *L() const { return el + EDGE_OFF; }
};
// edges of a tx block from the reconstruction (k_recon.cuh:750-775)
// fA / fL: the block's AvailU / AvailL for the edge filter type when they differ from the tx block's edges (a further
// 4:4:4 chroma tx block inside its block, k_recon.cuh:1654-1656); -1: the same
static void rc_edges(const Frame &F, PlaneEdges &E, int plane, int px, int py, int lw, int lh, int haveL, int haveA, int haveAR,
                     int haveBL, int mi_row, int mi_col, int fA = -1, int fL = -1) {
    const int sx = plane ? F.ssx : 0, sy = plane ? F.ssy : 0;
    E.maxX = ((F.MiCols * 4) >> sx) - 1; E.maxY = ((F.MiRows * 4) >> sy) - 1;
    build_edges(F.rec[plane].data(), F.stride[plane], px, py, lw, lh, haveL, haveA, haveAR, haveBL, E.maxX, E.maxY,
                E.ea + EDGE_OFF, E.el + EDGE_OFF, F.bd);
    E.dcv = dc_value(E.A(), E.L(), haveL, haveA, lw, lh, F.bd);
    E.haveL = haveL; E.haveA = haveA; E.x = px; E.y = py; E.lw = lw; E.lh = lh;
    E.ft = F.edgeFilter ? filter_type(plane, mi_row, mi_col, fA < 0 ? haveA : fA, fL < 0 ? haveL : fL, F.mY.data(), F.mUV.data(), F.MiCols,
                                      F.ssx, F.ssy) : 0;
}
static void pred_of(const Frame &F, const PlaneEdges &E, int c, int *pred) {
    predict_block(cand_mode(c), cand_delta(c), F.edgeFilter, E.ft, E.haveL, E.haveA, E.lw, E.lh, E.x, E.y, E.maxX, E.maxY,
                  E.A(), E.L(), E.dcv, pred, F.bd);
}

struct LumaOut { int mode, delta, eob; };
static LumaOut rc_luma(Frame &F, const Params &P, SbState &S, int sb, int bsize, int x, int y, int availU, int availL,
                       const TileBox &tb, float lambda) {
    const Tables &T = tables();
    const CostTab &C = T.cost[F.qctx];
    const int txSz = av1t::Max_Tx_Size_Rect[bsize], W = txw(txSz), H = txh(txSz);
    const int lw = av1t::Tx_Width_Log2[txSz], lh = av1t::Tx_Height_Log2[txSz];
    const int sbr4 = (y >> 2) & 15, sbc4 = (x >> 2) & 15, stepX = W >> 2, stepY = H >> 2;
    const int haveAR = S.bdm[0][sbr4 - 1 + 1][sbc4 + stepX + 1], haveBL = S.bdm[0][sbr4 + stepY + 1][sbc4 - 1 + 1];
    const int mi_row = y >> 2, mi_col = x >> 2;
    PlaneEdges E;
    rc_edges(F, E, 0, x, y, lw, lh, availL, availU, haveAR, haveBL, mi_row, mi_col);
    const BlockRes &br = F.blk[bsize][(size_t)(y / (4 * bh4(bsize))) * F.nbx[bsize] + x / (4 * bw4(bsize))];
    int skipCtx, dcCtx;
    txb_ctx(F, 0, txSz, x >> 2, y >> 2, bsize, tb, &skipCtx, &dcCtx);
    const int am = av1t::Intra_Mode_Context[availU ? F.mY[(size_t)(mi_row - 1) * F.MiCols + mi_col] : DC_PRED];
    const int lm = av1t::Intra_Mode_Context[availL ? F.mY[(size_t)mi_row * F.MiCols + mi_col - 1] : DC_PRED];
    auto mbits = [&](int c) -> float {
        const int m = cand_mode(c);
        float b = C.y_mode[am][lm][m];
        if (bsize >= BLOCK_8X8 && is_directional(m)) b += C.angle_delta[m - V_PRED][cand_delta(c) + 3];
        return b;
    };
    EvalIn e;
    e.txSz = txSz; e.plane = 0; e.dcQ = dc_q(F.qindex, F.bd); e.acQ = ac_q(F.qindex, F.bd); e.rnd = P.rnd; e.zbin = P.zbin;
    e.skipCtx = skipCtx; e.dcCtx = dcCtx; e.bd = F.bd; e.C = &C; e.resEnergy = 0.f;
    const int sp = F.stride[0];
    const uint16_t *src = F.src[0].data() + (size_t)y * sp + x;
    const int nc = std::max(1, std::min((int)br.ncand, P.Krecheck));
    const int set = tx_set_intra(txSz);
    // tasks: every re-checked candidate x every tx type of the set (txAllCL), or the candidates with DCT_DCT (+ the
    // analysis winner's tx type) followed by a tx-type pass for the winner (k_recon.cuh:943-1056)
    static thread_local int pred[64 * 64];
    static thread_local int16_t res[64 * 64], q[32 * 32], bestQ[32 * 32];
    int bestC = br.cand[0], bestT = br.tx, bestEob = -1;
    float bestJ = INF;
    auto run = [&](int c, int t) {
        pred_of(F, E, c, pred);
        for (int i = 0; i < H; i++) for (int j = 0; j < W; j++) res[i * W + j] = (int16_t)((int)src[i * sp + j] - pred[i * W + j]);
        EvalIn ee = e;
        ee.txType = t; ee.intraDir = cand_mode(c);
        ee.resEnergy = (W == 64 || H == 64) ? res_energy(res, W * H) : 0.f;
        const EvalOut o = eval_txb(res, ee, q);
        count_eval(F, 1, ee);
        const float J = fmaf(lambda, o.R + mbits(c), o.D);
        if (J < bestJ) { bestJ = J; bestC = c; bestT = t; bestEob = o.eob; memcpy(bestQ, q, sizeof(int16_t) * tx_coeffs(txSz)); }
    };
    if (P.txAllCL && set != TX_SET_DCTONLY) {
        for (int k = 0; k < nc; k++)
            for (int t = 0; t < 16; t++)
                if (tx_in_set(set, t)) run(br.cand[k], t);
    } else {
        const bool pass2 = P.txSearchCL && set != TX_SET_DCTONLY;
        int n = nc + (br.tx != DCT_DCT ? 1 : 0);
        if (n > 1 || pass2) {
            for (int k = 0; k < nc; k++) run(br.cand[k], DCT_DCT);
            if (br.tx != DCT_DCT) run(br.cand[0], br.tx);
            if (pass2) {
                const int c0 = bestC;
                for (int t = 1; t < 16; t++)
                    if (tx_in_set(set, t) && !(c0 == br.cand[0] && t == br.tx)) run(c0, t);
            }
        } else {
            run(br.cand[0], DCT_DCT);
        }
    }
    // final: the winner's levels, its prediction, exact inverse + reconstruction, record
    pred_of(F, E, bestC, pred);
    uint16_t *dst = F.rec[0].data() + (size_t)y * sp + x;
    for (int i = 0; i < H; i++) for (int j = 0; j < W; j++) dst[i * sp + j] = (uint16_t)pred[i * W + j];
    if (bestEob > 0) F.stats[3] += inverse_add(F, bestQ, txSz, bestT, e.dcQ, e.acQ, dst, sp) ? 1 : 0;
    const int eob = record_txb(F, S, sb, 0, txSz, bestT, x >> 2, y >> 2, bestQ);
    return LumaOut{cand_mode(bestC), cand_delta(bestC), eob};
}

struct ChromaOut { int uv, duv, aU, aV, eobU, eobV; };
static ChromaOut rc_chroma(Frame &F, const Params &P, SbState &S, int sb, int bsize, int mi_row, int mi_col, int yMode,
                           int availUC, int availLC, const TileBox &tb, float lambda) {
    const Tables &T = tables();
    const CostTab &C = T.cost[F.qctx];
    const int ssx = F.ssx, ssy = F.ssy, bd = F.bd;
    const int txSz = uv_tx_size(bsize, ssx, ssy), W = txw(txSz), H = txh(txSz);
    const int lw = av1t::Tx_Width_Log2[txSz], lh = av1t::Tx_Height_Log2[txSz];
    const int cx = (mi_col >> ssx) * 4, cy = (mi_row >> ssy) * 4, x4 = cx >> 2, y4 = cy >> 2;
    const int sbr4 = (mi_row & 15) >> ssy, sbc4 = (mi_col & 15) >> ssx, stepX = W >> 2, stepY = H >> 2;
    PlaneEdges E[3];
    for (int pl = 1; pl <= 2; pl++) {
        const int haveAR = S.bdm[pl][sbr4 - 1 + 1][sbc4 + stepX + 1], haveBL = S.bdm[pl][sbr4 + stepY + 1][sbc4 - 1 + 1];
        rc_edges(F, E[pl], pl, cx, cy, lw, lh, availLC, availUC, haveAR, haveBL, mi_row, mi_col);
    }
    const uint16_t *src[3] = {nullptr, F.src[1].data() + (size_t)cy * F.stride[1] + cx, F.src[2].data() + (size_t)cy * F.stride[2] + cx};
    const int sp = F.stride[1];
    const bool cflOk = P.cfl && cfl_allowed(bsize);
    const float lamS = P.satdLambda * T.lamS[bd == 10][F.qindex];   // the luma lambda's screening weight
    auto uvbits = [&](int m) -> float { return cfl_allowed(bsize) ? C.uv_cfl_yes[yMode][m] : C.uv_cfl_no[yMode][m]; };
    auto uvcbits = [&](int c) -> float {
        const int m = cand_mode(c);
        float b = uvbits(m);
        if (bsize >= BLOCK_8X8 && is_directional(m)) b += C.angle_delta[m - V_PRED][cand_delta(c) + 3];
        return b;
    };
    // ---- CfL luma values and alpha (k_recon.cuh:1254-1355) ----
    static thread_local int cflL[32 * 32];
    int cflAvg = 0, alpha[2] = {0, 0};
    if (cflOk) {
        const int lt = av1t::Max_Tx_Size_Rect[bsize], tw = txw(lt), th = txh(lt);   // tx_depth 0: one luma tx block
        int maxLumaW = 0, maxLumaH = 0;
        for (int ly = mi_row * 4; ly < mi_row * 4 + 4 * bh4(bsize); ly += th)
            for (int lx = mi_col * 4; lx < mi_col * 4 + 4 * bw4(bsize); lx += tw)
                if (lx < F.MiCols * 4 && ly < F.MiRows * 4) { maxLumaW = lx + tw; maxLumaH = ly + th; }
        const uint16_t *luma = F.rec[0].data();
        const int lp = F.stride[0];
        int sum = 0;
        for (int i = 0; i < H; i++)
            for (int j = 0; j < W; j++) {
                int v;
                if (ssx) {
                    int ly = (cy + i) << 1, lx = (cx + j) << 1;
                    if (ly > maxLumaH - 2) ly = maxLumaH - 2;
                    if (lx > maxLumaW - 2) lx = maxLumaW - 2;
                    v = (luma[ly * lp + lx] + luma[ly * lp + lx + 1] + luma[(ly + 1) * lp + lx] + luma[(ly + 1) * lp + lx + 1]) << 1;
                } else {
                    v = luma[std::min(cy + i, maxLumaH - 1) * lp + std::min(cx + j, maxLumaW - 1)] << 3;
                }
                cflL[i * W + j] = v;
                sum += v;
            }
        cflAvg = round2(sum, lw + lh);
        int64_t num[2] = {0, 0}, den = 0;
        for (int i = 0; i < H; i++)
            for (int j = 0; j < W; j++) {
                const int64_t l = cflL[i * W + j] - cflAvg;
                num[0] += l * ((int)src[1][i * sp + j] - E[1].dcv);
                num[1] += l * ((int)src[2][i * sp + j] - E[2].dcv);
                den += l * l;
            }
        for (int pl = 0; pl < 2; pl++) {
            const int a = den > 0 ? (int)rintf(64.f * (float)num[pl] / (float)den) : 0;
            alpha[pl] = std::max(-16, std::min(16, a));
        }
        for (int pl = 0; pl < 2; pl++) {   // refine alpha by +-1 on the pixel SSE (exact integer sums)
            int64_t sse[3] = {0, 0, 0};
            for (int i = 0; i < H; i++)
                for (int j = 0; j < W; j++) {
                    const int l = cflL[i * W + j] - cflAvg, s = src[pl + 1][i * sp + j];
                    for (int k = 0; k < 3; k++) {
                        const int a = std::max(-16, std::min(16, alpha[pl] + k - 1));
                        const int d = s - clip1(E[pl + 1].dcv + round2signed(a * l, 6), bd);
                        sse[k] += (int64_t)d * d;
                    }
                }
            int bk = 1;
            for (int k = 0; k < 3; k++) if (sse[k] < sse[bk]) bk = k;
            alpha[pl] = std::max(-16, std::min(16, alpha[pl] + bk - 1));
        }
    }
    // ---- SATD screening of the 13 UV modes (unfiltered edges), then deltas around the best directional one ----
    static thread_local int pred[32 * 32];
    int uvc[19];
    float satd[19];
    auto screen = [&](int c) -> float {
        int t = 0;
        for (int pl = 1; pl <= 2; pl++) {
            predict_block(cand_mode(c), cand_delta(c), 0, 0, E[pl].haveL, E[pl].haveA, lw, lh, cx, cy, E[pl].maxX, E[pl].maxY,
                          E[pl].A(), E[pl].L(), E[pl].dcv, pred, bd);
            t += satd_sum(src[pl], sp, pred, W, H);
        }
        return (float)t * 0.25f;
    };
    for (int m = 0; m < 13; m++) { uvc[m] = cand_pack(m, 0); satd[m] = screen(uvc[m]); }
    int nuv = 13;
    if (P.uvDeltas && bsize >= BLOCK_8X8) {
        int bm = -1;
        float bj = INF;
        for (int m = V_PRED; m <= D67_PRED; m++) {
            const float j = fmaf(lamS, uvcbits(cand_pack(m, 0)), satd[m]);
            if (bm < 0 || j < bj) { bj = j; bm = m; }
        }
        for (int k = 0; k < 6; k++) { uvc[13 + k] = cand_pack(bm, k < 3 ? -(k + 1) : k - 2); satd[13 + k] = screen(uvc[13 + k]); }
        nuv = 19;
    }
    // ---- top uvK (+ CfL) RD tasks for U and V ----
    struct Task { int c, t, pl; float J; int eob; std::vector<int16_t> q; };
    std::vector<Task> tasks;
    {
        float js[19];
        bool used[19] = {false};
        for (int i = 0; i < nuv; i++) js[i] = fmaf(lamS, uvcbits(uvc[i]), satd[i]);
        const int K = std::max(1, std::min(P.uvK, 6));
        for (int k = 0; k < K; k++) {
            int b = -1;
            for (int i = 0; i < nuv; i++) if (!used[i] && (b < 0 || js[i] < js[b])) b = i;
            if (b < 0) break;
            used[b] = true;
            for (int pl = 1; pl <= 2; pl++) tasks.push_back(Task{uvc[b], chroma_tx_type(txSz, cand_mode(uvc[b])), pl, 0.f, 0, {}});
        }
        if (cflOk && (alpha[0] || alpha[1]))
            for (int pl = 1; pl <= 2; pl++) tasks.push_back(Task{255, chroma_tx_type(txSz, UV_CFL_PRED), pl, 0.f, 0, {}});
    }
    int skipCtx[3], dcCtx[3];
    txb_ctx(F, 1, txSz, x4, y4, bsize, tb, &skipCtx[1], &dcCtx[1]);
    txb_ctx(F, 2, txSz, x4, y4, bsize, tb, &skipCtx[2], &dcCtx[2]);
    const int dcQ[3] = {0, dc_q(F.qindex, bd), dc_q(F.qindex, bd)}, acQ[3] = {0, ac_q(F.qindex, bd), ac_q(F.qindex, bd)};
    static thread_local int16_t res[32 * 32], q[32 * 32];
    auto chroma_pred = [&](int pl, int c, int *out) {
        if (c == 255) {
            for (int i = 0; i < H; i++)
                for (int j = 0; j < W; j++)
                    out[i * W + j] = clip1(E[pl].dcv + round2signed(alpha[pl - 1] * (cflL[i * W + j] - cflAvg), 6), bd);
        } else {
            pred_of(F, E[pl], c, out);
        }
    };
    for (Task &t : tasks) {
        chroma_pred(t.pl, t.c, pred);
        for (int i = 0; i < H; i++) for (int j = 0; j < W; j++) res[i * W + j] = (int16_t)((int)src[t.pl][i * sp + j] - pred[i * W + j]);
        EvalIn e;
        e.txSz = txSz; e.txType = t.t; e.plane = t.pl; e.dcQ = dcQ[t.pl]; e.acQ = acQ[t.pl]; e.rnd = P.rndC; e.zbin = P.zbinC;
        e.skipCtx = skipCtx[t.pl]; e.dcCtx = dcCtx[t.pl]; e.intraDir = 0; e.bd = bd; e.C = &C; e.resEnergy = 0.f;
        const EvalOut o = eval_txb(res, e, q);
        count_eval(F, 1, e);
        t.J = fmaf(lambda, o.R, o.D);
        t.eob = o.eob;
        t.q.assign(q, q + tx_coeffs(txSz));
    }
    // ---- pick (U + V + mode bits), k_recon.cuh:1561-1584 ----
    int bestUv = DC_PRED, bestUvC = cand_pack(DC_PRED, 0), bestT = -1;
    float bJ = INF;
    for (size_t t = 0; t + 1 < tasks.size(); t += 2) {
        const int c = tasks[t].c;
        float J = tasks[t].J + tasks[t + 1].J;
        int uv;
        if (c == 255) {
            uv = UV_CFL_PRED;
            const int aU = alpha[0], aV = alpha[1];
            const int sU = aU == 0 ? 0 : aU < 0 ? 1 : 2, sV = aV == 0 ? 0 : aV < 0 ? 1 : 2;
            float b = uvbits(UV_CFL_PRED) + C.cfl_sign[sU * 3 + sV - 1];
            if (sU) b += C.cfl_alpha[(sU - 1) * 3 + sV][abs(aU) - 1];
            if (sV) b += C.cfl_alpha[(sV - 1) * 3 + sU][abs(aV) - 1];
            J = fmaf(lambda, b, J);
        } else {
            uv = cand_mode(c);
            J = fmaf(lambda, uvcbits(c), J);
        }
        if (J < bJ) { bJ = J; bestUv = uv; bestUvC = c; bestT = (int)t; }
    }
    const bool cfl = bestUv == UV_CFL_PRED;
    const int uvTxType = chroma_tx_type(txSz, bestUv);
    int eobs[2] = {0, 0};
    // 4:4:4 blocks with a 64-sample side hold several chroma tx blocks: the decision above is the first one's; each
    // further one (raster order, per plane) gets its edges from the reconstruction, its own contexts and a closed-loop
    // final quantization with the chosen mode, and is not coded outside the coded area (k_recon.cuh:1636-1670)
    const int nTxX = ssx ? 1 : (4 * bw4(bsize)) / W, nTxY = ssy ? 1 : (4 * bh4(bsize)) / H;
    for (int pl = 1; pl <= 2; pl++)
        for (int ty = 0; ty < nTxY; ty++)
            for (int tx = 0; tx < nTxX; tx++) {
                const int ox = tx * W, oy = ty * H;
                if (x4 + (ox >> 2) >= F.MiCols || y4 + (oy >> 2) >= F.MiRows) continue;
                uint16_t *dst = F.rec[pl].data() + (size_t)(cy + oy) * F.stride[pl] + cx + ox;
                if (tx == 0 && ty == 0) {
                    Task &t = tasks[bestT + pl - 1];
                    chroma_pred(pl, cfl ? 255 : bestUvC, pred);
                    for (int i = 0; i < H; i++) for (int j = 0; j < W; j++) dst[i * F.stride[pl] + j] = (uint16_t)pred[i * W + j];
                    if (t.eob > 0) F.stats[3] += inverse_add(F, t.q.data(), txSz, uvTxType, dcQ[pl], acQ[pl], dst, F.stride[pl]) ? 1 : 0;
                    eobs[pl - 1] |= record_txb(F, S, sb, pl, txSz, uvTxType, x4, y4, t.q.data());
                    continue;
                }
                const int r4 = sbr4 + (oy >> 2), c4 = sbc4 + (ox >> 2);
                PlaneEdges Et;
                rc_edges(F, Et, pl, cx + ox, cy + oy, lw, lh, availLC || tx > 0, availUC || ty > 0, S.bdm[pl][r4 - 1 + 1][c4 + stepX + 1],
                         S.bdm[pl][r4 + stepY + 1][c4 - 1 + 1], mi_row, mi_col, availUC, availLC);
                int sc, dc;
                txb_ctx(F, pl, txSz, x4 + (ox >> 2), y4 + (oy >> 2), bsize, tb, &sc, &dc);
                pred_of(F, Et, bestUvC, pred);
                const uint16_t *so = src[pl] + (size_t)oy * sp + ox;
                for (int i = 0; i < H; i++) for (int j = 0; j < W; j++) res[i * W + j] = (int16_t)((int)so[i * sp + j] - pred[i * W + j]);
                EvalIn e;
                e.txSz = txSz; e.txType = uvTxType; e.plane = pl; e.dcQ = dcQ[pl]; e.acQ = acQ[pl]; e.rnd = P.rndC; e.zbin = P.zbinC;
                e.skipCtx = sc; e.dcCtx = dc; e.intraDir = 0; e.bd = bd; e.C = &C; e.resEnergy = 0.f;
                const EvalOut o = eval_txb(res, e, q);
                count_eval(F, 1, e);
                for (int i = 0; i < H; i++) for (int j = 0; j < W; j++) dst[i * F.stride[pl] + j] = (uint16_t)pred[i * W + j];
                if (o.eob > 0) F.stats[3] += inverse_add(F, q, txSz, uvTxType, dcQ[pl], acQ[pl], dst, F.stride[pl]) ? 1 : 0;
                eobs[pl - 1] |= record_txb(F, S, sb, pl, txSz, uvTxType, x4 + (ox >> 2), y4 + (oy >> 2), q);
            }
    return ChromaOut{bestUv, cfl ? 0 : cand_delta(bestUvC), cfl ? alpha[0] : 0, cfl ? alpha[1] : 0, eobs[0], eobs[1]};
}

void closed_loop_sb(Frame &F, const Params &P, int sb) {
    const Tables &T = tables();
    const int sbc = sb % F.sbCols, sbr = sb / F.sbCols, sbx = sbc * 64, sby = sbr * 64;
    const TileBox tb = tile_of(F, sbx, sby);
    const int miRowEnd = tb.y1 >> 2, miColEnd = tb.x1 >> 2;
    SbState S;
    // clear_block_decoded_flags (spec 7.3.3; encoder.cpp:358-372)
    for (int p = 0; p < 3; p++) {
        const int sx = p ? F.ssx : 0, sy = p ? F.ssy : 0;
        const int sbW4 = (miColEnd - sbc * 16) >> sx, sbH4 = (miRowEnd - sbr * 16) >> sy;
        for (int yy = -1; yy <= 16; yy++)
            for (int xx = -1; xx <= 16; xx++) {
                uint8_t v = 0;
                if (yy < 0 && xx < sbW4) v = 1;
                else if (xx < 0 && yy < sbH4) v = 1;
                if (yy == (16 >> sy) && xx == -1) v = 0;
                S.bdm[p][yy + 1][xx + 1] = v;
            }
    }
    S.nTxb = 0; S.nCoef = 0;
    F.sbBlocks[sb].clear(); F.sbBlockTxb[sb].clear(); F.sbTxb[sb].clear(); F.sbCoef[sb].clear();
    const float lambda = T.lambda[F.bd == 10][F.qindex] * P.lambdaMul;
    for (const Leaf &lf : F.leaves[sb]) {
        const int bsize = lf.bsize, x = sbx + lf.x4 * 4, y = sby + lf.y4 * 4, mi_row = y >> 2, mi_col = x >> 2;
        auto inside = [&](int r, int c) { return c >= (tb.x0 >> 2) && c < miColEnd && r >= (tb.y0 >> 2) && r < miRowEnd; };
        const int availU = inside(mi_row - 1, mi_col), availL = inside(mi_row, mi_col - 1);
        const int hasChroma = has_chroma(bsize, mi_row, mi_col, F.mono, F.ssx, F.ssy);
        int availUC = 0, availLC = 0;
        if (hasChroma) {
            availUC = (F.ssy && bh4(bsize) == 1) ? inside(mi_row - 2, mi_col) : availU;
            availLC = (F.ssx && bw4(bsize) == 1) ? inside(mi_row, mi_col - 2) : availL;
        }
        const uint32_t firstTxb = (uint32_t)S.nTxb;
        const LumaOut lo = rc_luma(F, P, S, sb, bsize, x, y, availU, availL, tb, lambda);
        ChromaOut co = {0, 0, 0, 0, 0, 0};
        if (hasChroma) co = rc_chroma(F, P, S, sb, bsize, mi_row, mi_col, lo.mode, availUC, availLC, tb, lambda * P.lambdaCMul);
        const int skip = lo.eob == 0 && co.eobU == 0 && co.eobV == 0;
        const int txSz = av1t::Max_Tx_Size_Rect[bsize];
        for (int r = mi_row; r < mi_row + bh4(bsize) && r < F.MiRows; r++)
            for (int c = mi_col; c < mi_col + bw4(bsize) && c < F.MiCols; c++) {
                const size_t o = (size_t)r * F.MiCols + c;
                F.mY[o] = (uint8_t)lo.mode;
                if (hasChroma) F.mUV[o] = (uint8_t)co.uv;
                F.mSkip[o] = (uint8_t)skip;
                F.mMiSize[o] = (uint8_t)bsize;
                F.mTx[o] = (uint8_t)txSz;
            }
        RecBlock b;
        b.mi_row = (uint16_t)mi_row; b.mi_col = (uint16_t)mi_col; b.bsize = (uint8_t)bsize; b.skip = (uint8_t)skip; b.tx_depth = 0;
        b.y_mode = (uint8_t)lo.mode;
        b.angle_delta_y = (int8_t)(bsize >= BLOCK_8X8 && is_directional(lo.mode) ? lo.delta : 0);
        b.uv_mode = (uint8_t)(hasChroma ? co.uv : 0);
        b.angle_delta_uv = (int8_t)(hasChroma && bsize >= BLOCK_8X8 && is_directional(co.uv) ? co.duv : 0);
        b.cfl_alpha_u = (int8_t)co.aU; b.cfl_alpha_v = (int8_t)co.aV;
        b.filter_intra = 0; b.cdef_idx = 0; b.sb_qindex = 0;
        F.sbBlocks[sb].push_back(b);
        F.sbBlockTxb[sb].push_back(firstTxb);
        if (skip) F.stats[1]++;
        if (hasChroma && co.uv == UV_CFL_PRED) F.stats[2]++;
        F.stats[0]++;
    }
}

// ------------------------------------------------------------------------------------------------------------
// pack (k_misc.cu:46-106, pipeline.cpp:878-904): tiles raster, SBs raster inside each tile
// ------------------------------------------------------------------------------------------------------------
void pack(Frame &F) {
    F.blocks.clear(); F.blockTxb.clear(); F.txb.clear(); F.coeffs.clear();
    for (int tr = 0; tr < F.tileRows; tr++)
        for (int tc = 0; tc < F.tileCols; tc++)
            for (int r = F.tileRowStart[tr]; r < F.tileRowStart[tr + 1]; r += 16)
                for (int c = F.tileColStart[tc]; c < F.tileColStart[tc + 1]; c += 16) {
                    const int sb = (r >> 4) * F.sbCols + (c >> 4);
                    const uint32_t t0 = (uint32_t)F.txb.size(), c0 = (uint32_t)F.coeffs.size();
                    for (size_t k = 0; k < F.sbBlocks[sb].size(); k++) {
                        F.blocks.push_back(F.sbBlocks[sb][k]);
                        F.blockTxb.push_back(t0 + F.sbBlockTxb[sb][k]);
                    }
                    for (RecTxb t : F.sbTxb[sb]) { t.coeff_off += c0; F.txb.push_back(t); }
                    F.coeffs.insert(F.coeffs.end(), F.sbCoef[sb].begin(), F.sbCoef[sb].end());
                }
    F.blockTxb.push_back((uint32_t)F.txb.size());
}

// the whole path; threads > 1 runs the analysis blocks and the SBs of each wavefront step in parallel
void encode_frame(Frame &F, const Params &P, int threads) {
    if (threads < 1) threads = 1;
    auto par = [&](int n, const std::function<void(int)> &fn) {
        if (threads == 1 || n < 2) { for (int i = 0; i < n; i++) fn(i); return; }
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        for (int t = 0; t < std::min(threads, n); t++)
            th.emplace_back([&]() { for (int i; (i = next.fetch_add(1)) < n;) fn(i); });
        for (auto &t : th) t.join();
    };
    for (int s = 0; s < 22; s++) {
        if (F.blk[s].empty() || !((P.shapeMask >> s) & 1)) continue;
        const int nbx = F.nbx[s];
        par(F.nby[s], [&](int by) { for (int bx = 0; bx < nbx; bx++) analyse_block(F, P, s, bx, by); });
    }
    par(F.sbCols * F.sbRows, [&](int sb) { dp_sb(F, P, sb); });
    // wavefront per tile: step = (c - c0) + 2 (r - r0)
    std::vector<std::vector<int>> steps;
    for (int tr = 0; tr < F.tileRows; tr++)
        for (int tc = 0; tc < F.tileCols; tc++) {
            const int r0 = F.tileRowStart[tr] >> 4, r1 = (F.tileRowStart[tr + 1] + 15) >> 4;
            const int c0 = F.tileColStart[tc] >> 4, c1 = (F.tileColStart[tc + 1] + 15) >> 4;
            for (int r = r0; r < r1; r++)
                for (int c = c0; c < c1; c++) {
                    const size_t s = (size_t)(c - c0) + 2 * (size_t)(r - r0);
                    if (steps.size() <= s) steps.resize(s + 1);
                    steps[s].push_back(r * F.sbCols + c);
                }
        }
    for (auto &st : steps) par((int)st.size(), [&](int i) { closed_loop_sb(F, P, st[i]); });
    pack(F);
}

}  // namespace orc
