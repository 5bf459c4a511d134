// Device building blocks of the search, executed by a "group" of 32 lanes (half of a wave64; two groups per wave):
// the float forward transform, dead-zone quantizer, distortion and exact coefficient rate of one transform block,
// intra edge preparation and SATD.
//
// Arithmetic contract (shared with the CPU oracle, bit for bit): every forward-transform output is one fmaf chain in
// ascending k; sums over coefficients run on 32 virtual lanes (lane l takes elements l, l + 32, ... in order) combined
// by the xor butterfly 16, 8, 4, 2, 1; J = fmaf(lambda, bits, D).  Which physical lane computes a chain does not
// change a value, so the passes spread their chains over the group freely.
#pragma once
#include "av1m_common.h"

namespace av1m {

constexpr int GW = 32;   // lanes per group

constexpr int tx_of(int W, int H) {
    return W == H ? (W == 4 ? TX_4X4 : W == 8 ? TX_8X8 : W == 16 ? TX_16X16 : W == 32 ? TX_32X32 : TX_64X64)
         : (W == 4 && H == 8) ? TX_4X8 : (W == 8 && H == 4) ? TX_8X4 : (W == 8 && H == 16) ? TX_8X16 : (W == 16 && H == 8) ? TX_16X8
         : (W == 16 && H == 32) ? TX_16X32 : (W == 32 && H == 16) ? TX_32X16 : (W == 32 && H == 64) ? TX_32X64
         : (W == 64 && H == 32) ? TX_64X32 : (W == 4 && H == 16) ? TX_4X16 : (W == 16 && H == 4) ? TX_16X4
         : (W == 8 && H == 32) ? TX_8X32 : (W == 32 && H == 8) ? TX_32X8 : (W == 16 && H == 64) ? TX_16X64 : TX_64X16;
}

// LDS view of a type: accesses through it compile to ds_* instructions, also inside functions the compiler keeps out
// of line (where a plain pointer would lose its address space and go through flat instructions)
template <typename T> using lds = __attribute__((address_space(3))) T;
// read-only global data seen through the constant address space: wave-uniform indices load through the scalar unit
template <typename T> using cst = __attribute__((address_space(4))) const T;
// the generic pointer to an LDS object (for library calls such as atomicAdd)
template <typename T> __device__ __forceinline__ T *gen(lds<T> *p) { return (T *)p; }
// p (the same value in every lane) as a wave-uniform constant-space pointer; pointers reach out-of-line functions in
// vector registers, so uniformity has to be restated for scalar loads
template <typename T> __device__ __forceinline__ cst<T> *uni(const T *p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (cst<T> *)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int glane() { return threadIdx.x & 31; }
__device__ __forceinline__ int gid() { return threadIdx.x >> 5; }
// the value of lane (this lane xor M) within the 32-lane group, every lane active: M = 1, 2 by DPP quad permutes
// ([1,0,3,2], [2,3,0,1]), M = 4, 8, 16 by ds_swizzle in bit-mask mode (and 0x1f, or 0, xor M)
template <int M>
__device__ __forceinline__ int lane_xor(int v) {
    static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16, "lane_xor: one lane bit of the 32-lane group");
    if constexpr (M == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    else if constexpr (M == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    else return __builtin_amdgcn_ds_swizzle(v, 0x1F | (M << 10));
}
template <int M>
__device__ __forceinline__ float lane_xorf(float v) { return __int_as_float(lane_xor<M>(__float_as_int(v))); }
template <int M>
__device__ __forceinline__ long long lane_xorl(long long v) {
    const unsigned lo = (unsigned)lane_xor<M>((int)(unsigned)(v & 0xFFFFFFFFll)), hi = (unsigned)lane_xor<M>((int)(v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
// orders the group's LDS accesses (one wave executes both groups in lockstep; this stops compiler reordering)
__device__ __forceinline__ void gsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// butterfly reductions over the 32 lanes of the group: partners lane xor 16, 8, 4, 2, 1 in that order, every lane ends
// with the result
__device__ __forceinline__ float gsum(float v) {
    v += lane_xorf<16>(v); v += lane_xorf<8>(v); v += lane_xorf<4>(v); v += lane_xorf<2>(v); v += lane_xorf<1>(v);
    return v;
}
__device__ __forceinline__ int gsumi(int v) {
    v += lane_xor<16>(v); v += lane_xor<8>(v); v += lane_xor<4>(v); v += lane_xor<2>(v); v += lane_xor<1>(v);
    return v;
}
__device__ __forceinline__ long long gsuml(long long v) {
    v += lane_xorl<16>(v); v += lane_xorl<8>(v); v += lane_xorl<4>(v); v += lane_xorl<2>(v); v += lane_xorl<1>(v);
    return v;
}
__device__ __forceinline__ int gmaxi(int v) {
    v = max(v, lane_xor<16>(v)); v = max(v, lane_xor<8>(v)); v = max(v, lane_xor<4>(v)); v = max(v, lane_xor<2>(v));
    v = max(v, lane_xor<1>(v));
    return v;
}
// index of the smallest (v, i) over the group, the lowest index among equal values; -1 if every v is INF
template <int M>
__device__ __forceinline__ void argmin_step(float &v, int &i) {
    const float ov = lane_xorf<M>(v);
    const int oi = lane_xor<M>(i);
    if (ov < v || (ov == v && oi < i)) { v = ov; i = oi; }
}
__device__ __forceinline__ int gargmin(float v, int i) {
    argmin_step<16>(v, i); argmin_step<8>(v, i); argmin_step<4>(v, i); argmin_step<2>(v, i); argmin_step<1>(v, i);
    return v < INF ? i : -1;
}

// inputs of one tx block evaluation; CP: pointer to the cost tables (global, or an LDS copy)
template <typename CP>
struct EvalInT {
    int txSz, txType, plane, dcQ, acQ, skipCtx, dcCtx, intraDir, bd;
    float rnd, zbin, resEnergy;
    CP C;
};
using EvalIn = EvalInT<const CostTab *>;
struct EvalOut { float D, R; int eob; };

__device__ __forceinline__ int txsz_ctx(int t) { return (av1t::Tx_Size_Sqr[t] + av1t::Tx_Size_Sqr_Up[t] + 1) >> 1; }

// txb_skip, luma tx type and eob position bits of a tx block of size TSZ
template <int TSZ, typename E>
__device__ float base_bits(const E &e, int eob) {
    const auto &C = *e.C;
    constexpr int t = TSZ, ctx = txsz_ctx_of(t);
    if (eob == 0) return C.txb_skip[ctx][e.skipCtx][1];
    float b = C.txb_skip[ctx][e.skipCtx][0];
    const int pt = e.plane > 0;
    if (e.plane == 0) {
        constexpr int set = tx_set_intra(t), sqr = av1t::Tx_Size_Sqr[t];
        if constexpr (set == TX_SET_INTRA_1) {
            int idx = 0;
#pragma unroll
            for (int i = 0; i < 7; i++) idx = av1t::Tx_Type_Intra_Inv_Set1[i] == e.txType ? i : idx;
            b += C.tx_set1[sqr][e.intraDir][idx];
        } else if constexpr (set == TX_SET_INTRA_2) {
            int idx = 0;
#pragma unroll
            for (int i = 0; i < 5; i++) idx = av1t::Tx_Type_Intra_Inv_Set2[i] == e.txType ? i : idx;
            b += C.tx_set2[sqr][e.intraDir][idx];
        }
    }
    constexpr int lw = av1t::Tx_Width_Log2[t], lh = av1t::Tx_Height_Log2[t];
    constexpr int ms = (lw < 5 ? lw : 5) + (lh < 5 ? lh : 5) - 4;
    const int eobPt = eob <= 2 ? eob : 2 + (31 - __clz(eob - 1));
    const int ec = tx_class(e.txType) == TX_CLASS_2D ? 0 : 1;
    const float p = ms == 0 ? C.eob_pt16[pt][ec][eobPt - 1] : ms == 1 ? C.eob_pt32[pt][ec][eobPt - 1]
                  : ms == 2 ? C.eob_pt64[pt][ec][eobPt - 1] : ms == 3 ? C.eob_pt128[pt][ec][eobPt - 1]
                  : ms == 4 ? C.eob_pt256[pt][ec][eobPt - 1] : ms == 5 ? C.eob_pt512[pt][eobPt - 1] : C.eob_pt1024[pt][eobPt - 1];
    b += p;
    if (eobPt >= 3) {
        const int rem = eob - ((1 << (eobPt - 2)) + 1);
        b += C.eob_extra[ctx][pt][eobPt - 3][(rem >> (eobPt - 3)) & 1];
        b += (float)(eobPt - 3);
    }
    return b;
}

// scan position of index c in a tx block of the adjusted size TW x TH: class VERT scans rows (raster order), HORIZ
// columns (Mrow / Mcol scans), 2D the default zig-zag scan dscan
template <int TW, int TH, int CLS, typename SP>
__device__ __forceinline__ int scan_pos(int c, SP dscan) {
    if constexpr (CLS == TX_CLASS_VERT) return c;
    else if constexpr (CLS == TX_CLASS_HORIZ) return (c % TH) * TW + c / TH;
    else return dscan[c];
}

// bits of the coefficient at scan position c (< eob) of a tx block of size TSZ and class CLS given the levels q
// (natural order, stride = adjusted width); cbOff: Coeff_Base_Ctx_Offset[TSZ] (5 x 5)
template <int TSZ, int CLS, typename QP, typename SP, typename OP, typename CT>
__device__ float coeff_bits(int c, int eob, QP q, SP dscan, OP cbOff, int pt, int dcCtx, const CT &C) {
    constexpr int ctx = txsz_ctx_of(TSZ), adj = av1t::Adjusted_Tx_Size[TSZ];
    constexpr int bwl = av1t::Tx_Width_Log2[adj], width = 1 << bwl, height = av1t::Tx_Height[adj];
    const int pos = scan_pos<width, height, CLS>(c, dscan), qv = q[pos], v = qv < 0 ? -qv : qv, level = v < 15 ? v : 15;
    const int row = pos >> bwl, col = pos - (row << bwl);
    float b;
    if (c == eob - 1) {
        const int bctx = c == 0 ? 0 : c <= (height << bwl) / 8 ? 1 : c <= (height << bwl) / 4 ? 2 : 3;
        b = C.coeff_base_eob[ctx][pt][bctx][min(level, 3) - 1];
    } else {
        int mag = 0;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const int rr = row + av1t::Sig_Ref_Diff_Offset[CLS][k][0], cc = col + av1t::Sig_Ref_Diff_Offset[CLS][k][1];
            if (rr < height && cc < width) { const int a = q[(rr << bwl) + cc]; mag += min(a < 0 ? -a : a, 3); }
        }
        int bctx = min((mag + 1) >> 1, 4);
        if constexpr (CLS == TX_CLASS_2D) bctx = (row == 0 && col == 0) ? 0 : bctx + cbOff[min(row, 4) * 5 + min(col, 4)];
        else bctx += min(CLS == TX_CLASS_VERT ? row : col, 2) == 0 ? av1t::Coeff_Base_Pos_Ctx_Offset[0]
                   : min(CLS == TX_CLASS_VERT ? row : col, 2) == 1 ? av1t::Coeff_Base_Pos_Ctx_Offset[1] : av1t::Coeff_Base_Pos_Ctx_Offset[2];
        b = C.coeff_base[ctx][pt][bctx][min(level, 3)];
    }
    if (level > 2) {
        int mag = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int rr = row + av1t::Mag_Ref_Offset_With_Tx_Class[CLS][k][0], cc = col + av1t::Mag_Ref_Offset_With_Tx_Class[CLS][k][1];
            if (rr < height && cc < width) { const int a = q[rr * width + cc]; mag += min(a < 0 ? -a : a, 15); }
        }
        mag = min((mag + 1) >> 1, 6);
        const int brctx = pos == 0 ? mag : CLS == TX_CLASS_2D ? ((row < 2 && col < 2) ? mag + 7 : mag + 14)
                        : CLS == TX_CLASS_HORIZ ? (col == 0 ? mag + 7 : mag + 14) : (row == 0 ? mag + 7 : mag + 14);
        const auto *br = C.coeff_br[min(ctx, (int)TX_32X32)][pt][brctx];
        int rem = level - 3;
        for (int i = 0; i < 4; i++) {
            const int k = min(rem, 3);
            b += br[k];
            rem -= k;
            if (k < 3) break;
        }
    }
    if (v) {
        b += c == 0 ? C.dc_sign[pt][dcCtx][qv < 0] : 1.0f;
        if (v > 14) b += (float)(2 * (31 - __clz(v - 14)) + 1);
    }
    return b;
}

// eob and rate of the levels q of a tx block of size TSZ, class CLS, by the group (all lanes return both); last: the
// lane's share of the eob (1 + the largest scan index of a nonzero level among the positions the lane quantized)
template <int TSZ, int CLS, typename QP, typename SP, typename OP, typename E>
__device__ void grp_rate(QP q, SP dscan, OP cbOff, const E &e, int last, int &eobOut, float &bitsOut) {
    const int lane = glane();
    const int eob = gmaxi(last);
    float bits = lane == 0 ? base_bits<TSZ>(e, eob) : 0.f;
    for (int c = lane; c < eob; c += GW) bits += coeff_bits<TSZ, CLS>(c, eob, q, dscan, cbOff, e.plane > 0, e.dcCtx, *e.C);
    eobOut = eob;
    bitsOut = gsum(bits);
}

// the default scan and the base-context offsets of tx size TSZ (global tables; the closed loop reads LDS copies)
template <int TSZ>
__device__ __forceinline__ const uint16_t *dscan_ptr(const DevTables &) { return scan_of(TSZ, DCT_DCT); }
template <int TSZ>
__device__ __forceinline__ const uint8_t *cboff_ptr(const DevTables &) { return &av1t::Coeff_Base_Ctx_Offset[TSZ][0][0]; }
template <int TSZ>
__device__ __forceinline__ const uint16_t *iscan_ptr(const DevTables &T) { return T.iscan + dscan_off(av1t::Adjusted_Tx_Size[TSZ]); }

// Evaluation scratch of one group in LDS, for tx blocks up to W x H
template <int W, int H>
struct EvalScr {
    static constexpr int TS = W + 1, TH = H > 32 ? 32 : H;
    int16_t res[W * H];
    float tmp[TH * TS];
    int16_t q[(W > 32 ? 32 : W) * TH];
};

template <int W, int H, typename FP, typename QP, typename E, typename TT>
__device__ __forceinline__ EvalOut grp_qrate(FP tmp, int ts, QP q, const E &e, const TT &T);

// Forward transform + quantization + distortion + rate of the residual res (W x H, row-major), by the 32 lanes of
// the calling group, with tmp (th x (W + 1) floats) as scratch.  Levels into q (natural order, tw x th).  All lanes
// return the same values.
template <int W, int H, typename RP, typename FP, typename QP, typename E, typename TT>
__device__ EvalOut grp_eval(RP res, FP tmp, QP q, const E &e, const TT &T) {
    constexpr int TW = W > 32 ? 32 : W, TH = H > 32 ? 32 : H, TS = W + 1;
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 4 ? 2 : H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    const int lane = glane();
    const int rk = row_kind(e.txType), ck = col_kind(e.txType), fud = flip_ud(e.txType), flr = flip_lr(e.txType);
    // Outputs per item chosen so that a pass fills the group's 32 lanes where the block allows (at most 8 per item):
    // each output is one fmaf chain over k ascending whichever lane computes it.
    // ---- column pass: item = (column j, OC consecutive outputs) ----
    {
        constexpr int OCF = W * TH / GW, OC = OCF < 1 ? 1 : OCF > 8 ? 8 : OCF, NI = W * (TH / OC);
        const auto *FcT = T.fwdT + T.fwdOff[ck][LH];
        const auto *Fc = T.fwd + T.fwdOff[ck][LH];
        for (int it = lane; it < NI; it += GW) {
            const int j = it % W, i0 = (it / W) * OC, jj = flr ? W - 1 - j : j;
            float acc[OC];
            if (ck == K_IDT) {
#pragma unroll
                for (int o = 0; o < OC; o++) acc[o] = Fc[(i0 + o) * H + i0 + o] * (float)res[(fud ? H - 1 - (i0 + o) : i0 + o) * W + jj];
            } else {
#pragma unroll
                for (int o = 0; o < OC; o++) acc[o] = 0.f;
#pragma unroll 4
                for (int k = 0; k < H; k++) {
                    const float x = (float)res[(fud ? H - 1 - k : k) * W + jj];
                    const auto *f = FcT + k * H + i0;
#pragma unroll
                    for (int o = 0; o < OC; o++) acc[o] = fmaf(f[o], x, acc[o]);
                }
            }
#pragma unroll
            for (int o = 0; o < OC; o++) tmp[(i0 + o) * TS + j] = acc[o];
        }
    }
    gsync();
    // ---- row pass, in place: item = (row i, OR consecutive outputs); every lane reads its row whole before any writes ----
    {
        constexpr int ORF = TW * TH / GW, OR = ORF < 1 ? 1 : ORF > TW ? TW : ORF, NR = TH * (TW / OR);
        static_assert(NR <= GW, "one row item per lane");
        const auto *FrT = T.fwdT + T.fwdOff[rk][LW];
        const auto *Fr = T.fwd + T.fwdOff[rk][LW];
        const float scale = T.txScale[e.txSz];
        const int i = lane % TH, o0 = (lane / TH) * OR;
        auto *row = tmp + i * TS;
        float acc[OR];
        if (lane < NR) {
            if (rk == K_IDT) {
#pragma unroll
                for (int o = 0; o < OR; o++) acc[o] = Fr[(o0 + o) * W + o0 + o] * row[o0 + o];
            } else {
#pragma unroll
                for (int o = 0; o < OR; o++) acc[o] = 0.f;
#pragma unroll 2
                for (int k = 0; k < W; k++) {
                    const float x = row[k];
                    const auto *f = FrT + k * W + o0;
#pragma unroll
                    for (int o = 0; o < OR; o++) acc[o] = fmaf(f[o], x, acc[o]);
                }
            }
        }
        gsync();
        if (lane < NR) {
#pragma unroll
            for (int o = 0; o < OR; o++) row[o0 + o] = acc[o] * scale;
        }
    }
    gsync();
    return grp_qrate<W, H>(tmp, TS, q, e, T);
}

// Quantization, distortion and rate of the coefficients of a W x H tx block held in tmp (row i of the kept tw x th
// region at tmp + i * ts), levels into q (natural order, tw x th).  All lanes of the group return the same values.
template <int W, int H, typename FP, typename QP, typename E, typename TT>
__device__ __forceinline__ EvalOut grp_qrate(FP tmp, int ts, QP q, const E &e, const TT &T) {
    constexpr int TW = W > 32 ? 32 : W, TH = H > 32 ? 32 : H;
    constexpr int LTW = TW == 4 ? 2 : TW == 8 ? 3 : TW == 16 ? 4 : 5;
    constexpr int N = TW * TH;
    const int lane = glane();
    constexpr int TSZ = tx_of(W, H);
    const int cls = tx_class(e.txType);
    // the scan of the rate: the default (2D) scan, rows (class VERT: raster order) or columns (class HORIZ)
    const int sk = (W > 16 || H > 16 || cls == TX_CLASS_2D) ? 0 : cls == TX_CLASS_HORIZ ? 1 : 2;
    const auto iscan = iscan_ptr<TSZ>(T);
    // ---- quantization + distortion (virtual lane = physical lane), and the eob ----
    const int den = dq_denom(e.txSz), dsh = den == 4 ? 2 : den == 2 ? 1 : 0;
    float dist = 0.f, kept = 0.f;
    int last = 0;
    for (int idx = lane; idx < N; idx += GW) {
        const float v = tmp[(idx >> LTW) * ts + (idx & (TW - 1))];
        const int step = idx == 0 ? e.dcQ : e.acQ;
        const float a = fabsf(v) * (float)den / (float)step;
        int l = a < e.zbin ? 0 : (int)(a + e.rnd);
        l = min(l, 16383);
        const int qv = v < 0 ? -l : l;
        q[idx] = (int16_t)qv;
        if (l) {
            const int c = sk == 0 ? (int)iscan[idx] : sk == 1 ? (idx & (TW - 1)) * TH + (idx >> LTW) : idx;
            last = max(last, c + 1);
        }
        const float dq = l ? (float)dequant(qv, step, dsh, e.bd) : 0.f;
        const float d = v - dq;
        const float sq = d * d;
        dist = dist + sq;
        kept = fmaf(v, v, kept);
    }
    const float distS = gsum(dist);
    gsync();
    const auto dscan = dscan_ptr<TSZ>(T);
    const auto cbOff = cboff_ptr<TSZ>(T);
    int eob;
    EvalOut o;
    if (sk == 0) grp_rate<TSZ, TX_CLASS_2D>(q, dscan, cbOff, e, last, eob, o.R);
    else if (sk == 1) grp_rate<TSZ, TX_CLASS_HORIZ>(q, dscan, cbOff, e, last, eob, o.R);
    else grp_rate<TSZ, TX_CLASS_VERT>(q, dscan, cbOff, e, last, eob, o.R);
    const float g = T.txg[e.txSz][e.txType];
    o.D = g * distS;
    if (W == 64 || H == 64) {
        const float keptS = gsum(kept);
        o.D = o.D + fmaxf(0.f, fmaf(-g, keptS, e.resEnergy));
    }
    o.eob = eob;
    return o;
}

// ---- the same evaluation by half a group (16 lanes) for tx blocks of up to 64 coefficients ----
// Lane l of the half holds virtual lanes l and l + 16 of the 32-lane form: their partial sums stay apart, and the sums
// over the group add them first (the xor-16 step of gsum), then run the remaining butterfly steps across the 16 lanes,
// so every result is the 32-lane form's to the bit.  The two halves of a group run two evaluations at once.
__device__ __forceinline__ int hlane() { return threadIdx.x & 15; }
__device__ __forceinline__ float hsum2(float a, float b) {
    float v = a + b;
    v += lane_xorf<8>(v); v += lane_xorf<4>(v); v += lane_xorf<2>(v); v += lane_xorf<1>(v);
    return v;
}
__device__ __forceinline__ int hmaxi(int v) {
    v = max(v, lane_xor<8>(v)); v = max(v, lane_xor<4>(v)); v = max(v, lane_xor<2>(v)); v = max(v, lane_xor<1>(v));
    return v;
}

template <int W, int H, typename FP, typename QP, typename E, typename TT>
__device__ EvalOut half_qrate(FP tmp, int ts, QP q, const E &e, const TT &T) {
    static_assert(W <= 16 && H <= 16 && W * H <= 64, "half-group evaluation: up to 64 coefficients");
    constexpr int TW = W, TH = H, LTW = TW == 4 ? 2 : TW == 8 ? 3 : 4, N = TW * TH, TSZ = tx_of(W, H);
    const int lane = hlane();
    const int cls = tx_class(e.txType);
    const int sk = cls == TX_CLASS_2D ? 0 : cls == TX_CLASS_HORIZ ? 1 : 2;
    const auto iscan = iscan_ptr<TSZ>(T);
    const int den = dq_denom(e.txSz), dsh = den == 4 ? 2 : den == 2 ? 1 : 0;
    float dist[2] = {0.f, 0.f};
    int last = 0;
#pragma unroll
    for (int h = 0; h < 2; h++)
        for (int idx = lane + 16 * h; idx < N; idx += GW) {
            const float v = tmp[(idx >> LTW) * ts + (idx & (TW - 1))];
            const int step = idx == 0 ? e.dcQ : e.acQ;
            const float a = fabsf(v) * (float)den / (float)step;
            int l = a < e.zbin ? 0 : (int)(a + e.rnd);
            l = min(l, 16383);
            const int qv = v < 0 ? -l : l;
            q[idx] = (int16_t)qv;
            if (l) {
                const int c = sk == 0 ? (int)iscan[idx] : sk == 1 ? (idx & (TW - 1)) * TH + (idx >> LTW) : idx;
                last = max(last, c + 1);
            }
            const float dq = l ? (float)dequant(qv, step, dsh, e.bd) : 0.f;
            const float d = v - dq;
            const float sq = d * d;
            dist[h] = dist[h] + sq;
        }
    const float distS = hsum2(dist[0], dist[1]);
    gsync();
    const auto dscan = dscan_ptr<TSZ>(T);
    const auto cbOff = cboff_ptr<TSZ>(T);
    const int eob = hmaxi(last);
    float b0 = lane == 0 ? base_bits<TSZ>(e, eob) : 0.f, b1 = 0.f;
    if (sk == 0) {
        for (int c = lane; c < eob; c += GW) b0 += coeff_bits<TSZ, TX_CLASS_2D>(c, eob, q, dscan, cbOff, e.plane > 0, e.dcCtx, *e.C);
        for (int c = lane + 16; c < eob; c += GW) b1 += coeff_bits<TSZ, TX_CLASS_2D>(c, eob, q, dscan, cbOff, e.plane > 0, e.dcCtx, *e.C);
    } else if (sk == 1) {
        for (int c = lane; c < eob; c += GW) b0 += coeff_bits<TSZ, TX_CLASS_HORIZ>(c, eob, q, dscan, cbOff, e.plane > 0, e.dcCtx, *e.C);
        for (int c = lane + 16; c < eob; c += GW) b1 += coeff_bits<TSZ, TX_CLASS_HORIZ>(c, eob, q, dscan, cbOff, e.plane > 0, e.dcCtx, *e.C);
    } else {
        for (int c = lane; c < eob; c += GW) b0 += coeff_bits<TSZ, TX_CLASS_VERT>(c, eob, q, dscan, cbOff, e.plane > 0, e.dcCtx, *e.C);
        for (int c = lane + 16; c < eob; c += GW) b1 += coeff_bits<TSZ, TX_CLASS_VERT>(c, eob, q, dscan, cbOff, e.plane > 0, e.dcCtx, *e.C);
    }
    EvalOut o;
    o.R = hsum2(b0, b1);
    o.D = T.txg[e.txSz][e.txType] * distS;
    o.eob = eob;
    return o;
}

// grp_eval by half a group (W x H up to 64 coefficients): the same column / row passes, items spread over 16 lanes
template <int W, int H, typename RP, typename FP, typename QP, typename E, typename TT>
__device__ EvalOut half_eval(RP res, FP tmp, QP q, const E &e, const TT &T) {
    constexpr int TW = W, TH = H, TS = W + 1;
    constexpr int LW = W == 4 ? 2 : W == 8 ? 3 : 4, LH = H == 4 ? 2 : H == 8 ? 3 : 4;
    const int lane = hlane();
    const int rk = row_kind(e.txType), ck = col_kind(e.txType), fud = flip_ud(e.txType), flr = flip_lr(e.txType);
    // outputs per item chosen so that each pass fills the 16 lanes (as grp_eval does with 32)
    {
        constexpr int OCF = W * TH / 16, OC = OCF < 1 ? 1 : OCF > 8 ? 8 : OCF, NI = W * (TH / OC);
        const auto *FcT = T.fwdT + T.fwdOff[ck][LH];
        const auto *Fc = T.fwd + T.fwdOff[ck][LH];
        for (int it = lane; it < NI; it += 16) {
            const int j = it % W, i0 = (it / W) * OC, jj = flr ? W - 1 - j : j;
            float acc[OC];
            if (ck == K_IDT) {
#pragma unroll
                for (int o = 0; o < OC; o++) acc[o] = Fc[(i0 + o) * H + i0 + o] * (float)res[(fud ? H - 1 - (i0 + o) : i0 + o) * W + jj];
            } else {
#pragma unroll
                for (int o = 0; o < OC; o++) acc[o] = 0.f;
#pragma unroll 4
                for (int k = 0; k < H; k++) {
                    const float x = (float)res[(fud ? H - 1 - k : k) * W + jj];
                    const auto *f = FcT + k * H + i0;
#pragma unroll
                    for (int o = 0; o < OC; o++) acc[o] = fmaf(f[o], x, acc[o]);
                }
            }
#pragma unroll
            for (int o = 0; o < OC; o++) tmp[(i0 + o) * TS + j] = acc[o];
        }
    }
    gsync();
    {
        constexpr int ORF = TW * TH / 16, OR = ORF < 1 ? 1 : ORF > TW ? TW : ORF, NR = TH * (TW / OR);
        static_assert(NR <= 16, "one row item per lane");
        const auto *FrT = T.fwdT + T.fwdOff[rk][LW];
        const auto *Fr = T.fwd + T.fwdOff[rk][LW];
        const float scale = T.txScale[e.txSz];
        const int i = lane % TH, o0 = (lane / TH) * OR;
        auto *row = tmp + i * TS;
        float acc[OR];
        if (lane < NR) {
            if (rk == K_IDT) {
#pragma unroll
                for (int o = 0; o < OR; o++) acc[o] = Fr[(o0 + o) * W + o0 + o] * row[o0 + o];
            } else {
#pragma unroll
                for (int o = 0; o < OR; o++) acc[o] = 0.f;
#pragma unroll 2
                for (int k = 0; k < W; k++) {
                    const float x = row[k];
                    const auto *f = FrT + k * W + o0;
#pragma unroll
                    for (int o = 0; o < OR; o++) acc[o] = fmaf(f[o], x, acc[o]);
                }
            }
        }
        gsync();   // every read of the rows precedes the in-place writes
        if (lane < NR) {
#pragma unroll
            for (int o = 0; o < OR; o++) row[o0 + o] = acc[o] * scale;
        }
    }
    gsync();
    return half_qrate<W, H>(tmp, TS, q, e, T);
}

// 2-D forward DCT_DCT of a W x H residual with a side of 32 or 64 by the 32 lanes of the group, the residual produced
// on the fly: res(k0, j, r) fills r[0 .. CH) with rows k0 .. k0 + CH - 1 of column j.  Column pass: lane = column (two columns, j and j + 32, for W = 64), one fmaf chain per kept
// output over k ascending with the matrix coefficient a wave-uniform scalar operand; row pass: lane = row.  The
// coefficients (scaled) land in tmp (row i at tmp + i * (W + 1), tw x th kept).  Returns the residual energy sum over the
// raster order (W or H = 64: the energy the dropped coefficients carry), else 0.
template <int W, int H, typename Res>
__device__ __forceinline__ float grp_fwd_dct(lds<float> *tmp, const DevTables &T, float scale, Res res) {
    constexpr int TW = W > 32 ? 32 : W, TH = H > 32 ? 32 : H, TS = W + 1;
    constexpr int LW = W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6, LH = H == 8 ? 3 : H == 16 ? 4 : H == 32 ? 5 : 6;
    constexpr bool E64 = W == 64 || H == 64;
    constexpr int NC = W == 64 ? 2 : 1;
    constexpr int CH = H < 8 ? H : 8;   // residual rows made per call of res (its mode tests once per run)
    const int lane = glane();
    cst<DevTables> *TU = uni(&T);
    cst<float> *FcT = TU->fwdT + TU->fwdOff[K_DCT][LH];
    cst<float> *FrT = TU->fwdT + TU->fwdOff[K_DCT][LW];
    float en = 0.f, eo = 0.f;
    if (lane < W) {
        float acc[NC][TH];
#pragma unroll
        for (int i = 0; i < TH; i++) { acc[0][i] = 0.f; if constexpr (NC == 2) acc[NC - 1][i] = 0.f; }
        for (int k0 = 0; k0 < H; k0 += CH) {
            int ra[CH], rb[CH];
            res(k0, lane, ra);
            if constexpr (NC == 2) res(k0, lane + 32, rb);
#pragma unroll
            for (int kk = 0; kk < CH; kk++) {
                const int k = k0 + kk, r0 = ra[kk], r1 = NC == 2 ? rb[kk] : 0;
                if constexpr (E64) {
                    // raster-order virtual lane l holds elements l, l + 32, ...: W = 64 alternates columns l and l + 32
                    // per row; W = 32 is column l; W = 16 (16x64) splits column l & 15 into even (lane l) and odd rows
                    // (l + 16)
                    if constexpr (W == 64) { en = en + (float)(r0 * r0); en = en + (float)(r1 * r1); }
                    else if constexpr (W == 32) en = en + (float)(r0 * r0);
                    else { if (k & 1) eo = eo + (float)(r0 * r0); else en = en + (float)(r0 * r0); }
                }
                const float x0 = (float)r0, x1 = (float)r1;
                cst<float> *f = FcT + k * H;
#pragma unroll
                for (int i = 0; i < TH; i++) {
                    const float fi = f[i];
                    acc[0][i] = fmaf(fi, x0, acc[0][i]);
                    if constexpr (NC == 2) acc[NC - 1][i] = fmaf(fi, x1, acc[NC - 1][i]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < TH; i++) {
            tmp[i * TS + lane] = acc[0][i];
            if constexpr (NC == 2) tmp[i * TS + lane + 32] = acc[NC - 1][i];
        }
    }
    if constexpr (E64 && W == 16) {
        const float o = __shfl_xor(eo, 16, 32);
        if (lane >= 16) en = o;
    }
    gsync();
    if (lane < TH) {
        float acc[TW];
#pragma unroll
        for (int o = 0; o < TW; o++) acc[o] = 0.f;
        lds<float> *row = tmp + lane * TS;
#pragma unroll 2
        for (int k = 0; k < W; k++) {
            const float x = row[k];
            cst<float> *f = FrT + k * W;
#pragma unroll
            for (int o = 0; o < TW; o++) acc[o] = fmaf(f[o], x, acc[o]);
        }
#pragma unroll
        for (int o = 0; o < TW; o++) row[o] = acc[o] * scale;
    }
    gsync();
    return E64 ? gsum(en) : 0.f;
}

// sum of r^2 over a W x H residual in raster order (32 virtual lanes)
template <int W, int H, typename RP>
__device__ float grp_res_energy(RP res) {
    float e = 0.f;
    for (int p = glane(); p < W * H; p += GW) e += (float)(res[p] * res[p]);
    return gsum(e);
}

// ---- intra edge preparation by a group: av1g_edge_prepare's integer arithmetic (spec 7.11.2.4 step 4, 7.11.2.9-12)
// buf: AboveRow or LeftCol, index 0 = first sample (indices -2 .. 2 (w + h) used)
template <typename P>
__device__ void grp_edge_filter(P *buf, int sz, int strength) {
    if (!strength) return;
    const int lane = glane();
    const uint8_t *K = av1t::Intra_Edge_Kernel[strength - 1];
    int v[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int i = lane + 1 + GW * k;
        int s = 0;
        if (i < sz) {
#pragma unroll
            for (int j = 0; j < 5; j++) s += K[j] * buf[clip3(0, sz - 1, i - 2 + j) - 1];
        }
        v[k] = (s + 8) >> 4;
    }
    gsync();
#pragma unroll
    for (int k = 0; k < 5; k++)
        if (lane + 1 + GW * k < sz) buf[lane + GW * k] = v[k];
    gsync();
}
template <typename P>
__device__ void grp_edge_upsample(P *buf, int numPx, int bd) {   // numPx <= 16
    const int lane = glane();
    auto dup = [&](int k) -> int { return k == 0 ? buf[-1] : k == numPx + 2 ? buf[numPx - 1] : buf[k - 2]; };
    int s = 0, e = 0;
    if (lane < numPx) {
        s = clip3(0, (1 << bd) - 1, round2(-dup(lane) + 9 * dup(lane + 1) + 9 * dup(lane + 2) - dup(lane + 3), 4));
        e = dup(lane + 2);
    }
    const int c = buf[-1];
    gsync();
    if (lane < numPx) { buf[2 * lane - 1] = s; buf[2 * lane] = e; }
    if (lane == 0) buf[-2] = c;
    gsync();
}
// prepare directional-mode edges in place (A, L are the group's copies); returns the upsampling flags
template <typename P>
__device__ void grp_edge_prepare(int pa, int lw, int lh, int ft, int haveL, int haveA, int x, int y, int maxX, int maxY, P *A, P *L,
                                 int bd, int &upA, int &upL) {
    const int w = 1 << lw, h = 1 << lh;
    if (pa != 90 && pa != 180) {
        if (pa > 90 && pa < 180 && w + h >= 24) {
            if (glane() == 0) {
                const int c = round2(L[0] * 5 + A[-1] * 6 + A[0] * 5, 4);
                L[-1] = c;
                A[-1] = c;
            }
            gsync();
        }
        if (haveA) grp_edge_filter(A, min(w, maxX - x + 1) + (pa < 90 ? h : 0) + 1, edge_strength(w, h, ft, pa - 90));
        if (haveL) grp_edge_filter(L, min(h, maxY - y + 1) + (pa > 180 ? w : 0) + 1, edge_strength(w, h, ft, pa - 180));
    }
    upA = edge_upsample(w, h, ft, pa - 90);
    if (upA) grp_edge_upsample(A, w + (pa < 90 ? h : 0), bd);
    upL = edge_upsample(w, h, ft, pa - 180);
    if (upL) grp_edge_upsample(L, h + (pa > 180 ? w : 0), bd);
}

// build AboveRow / LeftCol [-1 .. w + h - 1] of the tx block at (x, y) of plane (pitch) by the group's lanes
template <typename SP, typename P>
__device__ void grp_build_edges(const SP *plane, int pitch, int x, int y, int lw, int lh, int haveL, int haveA, int haveAR, int haveBL,
                                int maxX, int maxY, P *A, P *L, int bd) {
    const int w = 1 << lw, h = 1 << lh, n = w + h;
    const int alim = min(x + (haveAR ? 2 * w : w) - 1, maxX), llim = min(y + (haveBL ? 2 * h : h) - 1, maxY);
    for (int i = glane(); i < n; i += GW) {
        int a, l;
        if (!haveA && haveL) a = plane[(size_t)y * pitch + x - 1];
        else if (!haveA) a = (1 << (bd - 1)) - 1;
        else a = plane[(size_t)(y - 1) * pitch + min(x + i, alim)];
        if (!haveL && haveA) l = plane[(size_t)(y - 1) * pitch + x];
        else if (!haveL) l = (1 << (bd - 1)) + 1;
        else l = plane[(size_t)min(y + i, llim) * pitch + x - 1];
        A[i] = a;
        L[i] = l;
    }
    if (glane() == 0) {
        int c;
        if (haveA && haveL) c = plane[(size_t)(y - 1) * pitch + x - 1];
        else if (haveA) c = plane[(size_t)(y - 1) * pitch + x];
        else if (haveL) c = plane[(size_t)y * pitch + x - 1];
        else c = 1 << (bd - 1);
        A[-1] = c;
        L[-1] = c;
    }
    gsync();
}
// DC value (spec 7.11.2.5) with the edge sums over the group
template <typename P>
__device__ int grp_dc_value(const P *A, const P *L, int haveL, int haveA, int lw, int lh, int bd) {
    const int w = 1 << lw, h = 1 << lh;
    int sa = 0, sl = 0;
    for (int k = glane(); k < w; k += GW) sa += A[k];
    for (int k = glane(); k < h; k += GW) sl += L[k];
    sa = gsumi(sa);
    sl = gsumi(sl);
    if (haveL && haveA) return (sa + sl + ((w + h) >> 1)) / (w + h);
    if (haveL) return clip3(0, (1 << bd) - 1, (sl + (h >> 1)) >> lh);
    if (haveA) return clip3(0, (1 << bd) - 1, (sa + (w >> 1)) >> lw);
    return 1 << (bd - 1);
}

// sum |4x4 Hadamard coefficients| of d
__device__ __forceinline__ int hadamard4(int *d) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int *r = d + 4 * i;
        const int s0 = r[0] + r[1], s1 = r[0] - r[1], s2 = r[2] + r[3], s3 = r[2] - r[3];
        r[0] = s0 + s2; r[1] = s1 + s3; r[2] = s0 - s2; r[3] = s1 - s3;
    }
    int sum = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int s0 = d[j] + d[4 + j], s1 = d[j] - d[4 + j], s2 = d[8 + j] + d[12 + j], s3 = d[8 + j] - d[12 + j];
        sum += abs(s0 + s2) + abs(s1 + s3) + abs(s0 - s2) + abs(s1 - s3);
    }
    return sum;
}

// sum over the SEG consecutive lanes of a segment (SEG a power of 2 up to 32), in every lane of it
template <int SEG>
__device__ __forceinline__ int seg_sumi(int v) {
    if constexpr (SEG >= 32) v += lane_xor<16>(v);
    if constexpr (SEG >= 16) v += lane_xor<8>(v);
    if constexpr (SEG >= 8) v += lane_xor<4>(v);
    if constexpr (SEG >= 4) v += lane_xor<2>(v);
    if constexpr (SEG >= 2) v += lane_xor<1>(v);
    return v;
}

// the same transform with one value of d per lane: lanes 16k .. 16k + 15 of the group hold one 4x4 block in raster
// order (lane bits 0-1: column, 2-3: row).  The butterflies are those of hadamard4 (rows, then columns), so the 16
// lanes return the absolute values of its 16 coefficients, whose sum is its result.  Every lane of the group calls it.
__device__ __forceinline__ int hadamard4_lanes(int v) {
    const int l = glane();
    int t = lane_xor<1>(v);
    v = (l & 1) ? t - v : v + t;
    t = lane_xor<2>(v);
    v = (l & 2) ? t - v : v + t;
    t = lane_xor<4>(v);
    v = (l & 4) ? t - v : v + t;
    t = lane_xor<8>(v);
    v = (l & 8) ? t - v : v + t;
    return abs(v);
}

// runtime tx-size dispatch to a template taking <W, H>
#define AV1M_TX_SWITCH(txSz, M)                                                                                   \
    switch (txSz) {                                                                                               \
    case TX_4X4: M(4, 4); break; case TX_8X8: M(8, 8); break; case TX_16X16: M(16, 16); break;                    \
    case TX_32X32: M(32, 32); break; case TX_64X64: M(64, 64); break; case TX_4X8: M(4, 8); break;                \
    case TX_8X4: M(8, 4); break; case TX_8X16: M(8, 16); break; case TX_16X8: M(16, 8); break;                    \
    case TX_16X32: M(16, 32); break; case TX_32X16: M(32, 16); break; case TX_32X64: M(32, 64); break;            \
    case TX_64X32: M(64, 32); break; case TX_4X16: M(4, 16); break; case TX_16X4: M(16, 4); break;                \
    case TX_8X32: M(8, 32); break; case TX_32X8: M(32, 8); break; case TX_16X64: M(16, 64); break;                \
    default: M(64, 16); break;                                                                                    \
    }

}  // namespace av1m
