// Shared host/device definitions of the MI355X intra RD core: AV1 symbol values, block / transform geometry, the exact
// dequantizer and inverse transforms, intra prediction primitives, and the structures the kernels exchange.
// Compiled by hipcc for gfx950 device code and for the host side of the library.
#pragma once
#include <stdint.h>
#include "av1_spec_tables.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AV1M_HD __host__ __device__ __forceinline__
#else
#define AV1M_HD inline
#endif

namespace av1m {

// ---- AV1 symbol values (spec section 6.10.x; the values of the reference's enums_gen.h) ----
enum { BLOCK_4X4 = 0, BLOCK_4X8, BLOCK_8X4, BLOCK_8X8, BLOCK_8X16, BLOCK_16X8, BLOCK_16X16, BLOCK_16X32, BLOCK_32X16,
       BLOCK_32X32, BLOCK_32X64, BLOCK_64X32, BLOCK_64X64, BLOCK_4X16 = 16, BLOCK_16X4, BLOCK_8X32, BLOCK_32X8,
       BLOCK_16X64, BLOCK_64X16, BLOCK_INVALID, NSHAPES = 22 };
enum { TX_4X4 = 0, TX_8X8, TX_16X16, TX_32X32, TX_64X64, TX_4X8, TX_8X4, TX_8X16, TX_16X8, TX_16X32, TX_32X16,
       TX_32X64, TX_64X32, TX_4X16, TX_16X4, TX_8X32, TX_32X8, TX_16X64, TX_64X16 };
enum { DCT_DCT = 0, ADST_DCT, DCT_ADST, ADST_ADST, FLIPADST_DCT, DCT_FLIPADST, FLIPADST_FLIPADST, ADST_FLIPADST,
       FLIPADST_ADST, IDTX, V_DCT, H_DCT, V_ADST, H_ADST, V_FLIPADST, H_FLIPADST };
enum { DC_PRED = 0, V_PRED, H_PRED, D45_PRED, D135_PRED, D113_PRED, D157_PRED, D203_PRED, D67_PRED, SMOOTH_PRED,
       SMOOTH_V_PRED, SMOOTH_H_PRED, PAETH_PRED, UV_CFL_PRED };
enum { PARTITION_NONE = 0, PARTITION_HORZ, PARTITION_VERT, PARTITION_SPLIT, PARTITION_HORZ_A, PARTITION_HORZ_B,
       PARTITION_VERT_A, PARTITION_VERT_B, PARTITION_HORZ_4, PARTITION_VERT_4 };
enum { TX_SET_DCTONLY = 0, TX_SET_INTRA_1 = 1, TX_SET_INTRA_2 = 2 };
enum { TX_CLASS_2D = 0, TX_CLASS_HORIZ = 1, TX_CLASS_VERT = 2 };
enum { K_DCT = 0, K_ADST = 1, K_IDT = 2 };
constexpr int EDGE_OFF = 16, EDGE_LEN = 16 + 2 * 64 + 16;
constexpr float INF = 3.0e38f;
constexpr int MAXK = 6;
constexpr int MAX_LEAVES = 256;          // blocks per 64x64 superblock at most
constexpr int SB_TXB_CAP = 384;          // tx blocks per SB (4:2:0): 256 luma + 2 x 64 chroma
constexpr int SB_COEF_CAP = 6144;        // coefficients per SB (4:2:0): 64 x 64 + 2 x 32 x 32
constexpr int SB_TXB_CAP_444 = 768;      // 4:4:4: 3 x 256
constexpr int SB_COEF_CAP_444 = 12288;   // 4:4:4: 3 x 64 x 64
constexpr int FWD_FLOATS = (16 + 64 + 256 + 1024 + 4096) + (16 + 64 + 256) + (16 + 64 + 256 + 1024);

// Geometry of the ten AV1 partitions of an n x n node (spec 5.11.4 decode_partition): up to four parts in decode order,
// each the partition's sub-size (split = 0) or the node's split square (split = 1), placed at (dx, dy) quarters of n.
// Shared by the partition DP (device) and the partition feasibility check (host).
struct PartPlace { uint8_t split, dx, dy; };
struct PartShape { uint8_t parts; PartPlace at[4]; };
inline constexpr PartShape Part_Shape[10] = {
    {1, {{0, 0, 0}}},                                      // NONE
    {2, {{0, 0, 0}, {0, 0, 2}}},                           // HORZ
    {2, {{0, 0, 0}, {0, 2, 0}}},                           // VERT
    {4, {{1, 0, 0}, {1, 2, 0}, {1, 0, 2}, {1, 2, 2}}},     // SPLIT
    {3, {{1, 0, 0}, {1, 2, 0}, {0, 0, 2}}},                // HORZ_A
    {3, {{0, 0, 0}, {1, 0, 2}, {1, 2, 2}}},                // HORZ_B
    {3, {{1, 0, 0}, {1, 0, 2}, {0, 2, 0}}},                // VERT_A
    {3, {{0, 0, 0}, {1, 2, 0}, {1, 2, 2}}},                // VERT_B
    {4, {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 0, 3}}},     // HORZ_4
    {4, {{0, 0, 0}, {0, 1, 0}, {0, 2, 0}, {0, 3, 0}}},     // VERT_4
};
// the order in which the DP weighs the partitions of a node inside the frame (the first of equal costs is kept)
inline constexpr uint8_t Part_Order[10] = {PARTITION_SPLIT, PARTITION_NONE, PARTITION_HORZ, PARTITION_VERT, PARTITION_HORZ_A,
                                           PARTITION_HORZ_B, PARTITION_VERT_A, PARTITION_VERT_B, PARTITION_HORZ_4, PARTITION_VERT_4};

AV1M_HD int clip3(int lo, int hi, int x) { return x < lo ? lo : x > hi ? hi : x; }
AV1M_HD int round2(int x, int n) { return n == 0 ? x : (x + (1 << (n - 1))) >> n; }
AV1M_HD int64_t round2l(int64_t x, int n) { return n == 0 ? x : (x + ((int64_t)1 << (n - 1))) >> n; }
AV1M_HD int round2signed(int x, int n) { return x >= 0 ? round2(x, n) : -round2(-x, n); }
AV1M_HD int bw4(int bs) { return av1t::Num_4x4_Blocks_Wide[bs]; }
AV1M_HD int square_bsize(int n) { return n == 8 ? BLOCK_8X8 : n == 16 ? BLOCK_16X16 : n == 32 ? BLOCK_32X32 : BLOCK_64X64; }
// block size of part k of partition p of an n x n node (n >= 8)
AV1M_HD int part_bsize(int p, int k, int n) {
    return av1t::Partition_Subsize[Part_Shape[p].at[k].split ? PARTITION_SPLIT : p][square_bsize(n)];
}
AV1M_HD int bh4(int bs) { return av1t::Num_4x4_Blocks_High[bs]; }
AV1M_HD int tx_w(int t) { return av1t::Tx_Width[t]; }
AV1M_HD int tx_h(int t) { return av1t::Tx_Height[t]; }
AV1M_HD int tx_ncoef(int t) { return (tx_w(t) > 32 ? 32 : tx_w(t)) * (tx_h(t) > 32 ? 32 : tx_h(t)); }
AV1M_HD int is_dir(int m) { return m >= V_PRED && m <= D67_PRED; }
AV1M_HD int cand_mode(int c) { return c >> 3; }
AV1M_HD int cand_delta(int c) { return (c & 7) - 3; }
AV1M_HD int cand_pack(int m, int d) { return (m << 3) | (d + 3); }
AV1M_HD int pangle(int mode, int delta) { return is_dir(mode) ? av1t::Mode_To_Angle[mode] + 3 * delta : 0; }

AV1M_HD constexpr int tx_set_intra(int t) {
    if (av1t::Tx_Size_Sqr_Up[t] >= TX_32X32) return TX_SET_DCTONLY;
    return av1t::Tx_Size_Sqr[t] == TX_16X16 ? TX_SET_INTRA_2 : TX_SET_INTRA_1;
}
AV1M_HD int tx_in_set(int set, int t) { return av1t::Tx_Type_In_Set_Intra[set][t]; }
AV1M_HD int chroma_tx_type(int uvTx, int uvMode) {
    if (av1t::Tx_Size_Sqr_Up[uvTx] > TX_32X32) return DCT_DCT;
    const int t = av1t::Mode_To_Txfm[uvMode];
    return tx_in_set(tx_set_intra(uvTx), t) ? t : DCT_DCT;
}
AV1M_HD constexpr int tx_class(int t) {
    return (t == V_DCT || t == V_ADST || t == V_FLIPADST) ? TX_CLASS_VERT : (t == H_DCT || t == H_ADST || t == H_FLIPADST) ? TX_CLASS_HORIZ : TX_CLASS_2D;
}
// 1-D kinds of the row (horizontal) and column (vertical) passes; flips
AV1M_HD constexpr int row_kind(int t) {
    return (t == DCT_DCT || t == ADST_DCT || t == FLIPADST_DCT || t == H_DCT) ? K_DCT
         : (t == IDTX || t == V_DCT || t == V_ADST || t == V_FLIPADST) ? K_IDT : K_ADST;
}
AV1M_HD constexpr int col_kind(int t) {
    return (t == DCT_DCT || t == DCT_ADST || t == DCT_FLIPADST || t == V_DCT) ? K_DCT
         : (t == IDTX || t == H_DCT || t == H_ADST || t == H_FLIPADST) ? K_IDT : K_ADST;
}
// Algorithmic forward-transform FLOPs of one evaluation of a W x H transform block of type t (work accounting): 2 per
// fmaf of the column pass (W columns x min(H, 32) kept outputs x H taps) and of the row pass (min(H, 32) rows x
// min(W, 32) kept outputs x W taps), 1 per multiply of an identity kind's pass and of the row output scale.
AV1M_HD constexpr uint32_t fwd_flops(int W, int H, int t) {
    const uint32_t tw = W > 32 ? 32 : W, th = H > 32 ? 32 : H;
    const uint32_t col = col_kind(t) == K_IDT ? (uint32_t)W * th : 2u * W * th * H;
    const uint32_t row = row_kind(t) == K_IDT ? th * tw : 2u * th * tw * W;
    return col + row + th * tw;
}
// the same summed over the types of transform set `set` other than DCT_DCT
AV1M_HD constexpr uint32_t set_flops(int W, int H, int set) {
    uint32_t f = 0;
    for (int t = 1; t < 16; t++)
        if (av1t::Tx_Type_In_Set_Intra[set][t]) f += fwd_flops(W, H, t);
    return f;
}
AV1M_HD int flip_ud(int t) { return t == FLIPADST_DCT || t == FLIPADST_ADST || t == V_FLIPADST || t == FLIPADST_FLIPADST; }
AV1M_HD int flip_lr(int t) { return t == DCT_FLIPADST || t == ADST_FLIPADST || t == H_FLIPADST || t == FLIPADST_FLIPADST; }
AV1M_HD int dq_denom(int t) {
    return (t == TX_64X64 || t == TX_32X64 || t == TX_64X32) ? 4
         : (t == TX_32X32 || t == TX_16X32 || t == TX_32X16 || t == TX_16X64 || t == TX_64X16) ? 2 : 1;
}
AV1M_HD int dc_q(int q, int bd) { return av1t::Dc_Qlookup[(bd - 8) >> 1][clip3(0, 255, q)]; }
AV1M_HD int ac_q(int q, int bd) { return av1t::Ac_Qlookup[(bd - 8) >> 1][clip3(0, 255, q)]; }
// spec 7.12.3 step 1 without quantizer matrices: |level * q| & 0xFFFFFF, / dqDenom (a shift: 1, 2, 4), clamp 2^(7+bd)
AV1M_HD int dequant(int level, int q, int denShift, int bd) {
    const int dq = level * q;
    int a = ((dq < 0 ? -dq : dq) & 0xFFFFFF) >> denShift;
    if (dq < 0) a = -a;
    return clip3(-(1 << (7 + bd)), (1 << (7 + bd)) - 1, a);
}

AV1M_HD const uint16_t *scan_of(int t, int txType) {
    using namespace av1t;
    if (t == TX_16X64) return Default_Scan_16x32;
    if (t == TX_64X16) return Default_Scan_32x16;
    if (Tx_Size_Sqr_Up[t] == TX_64X64) return Default_Scan_32x32;
    const int cls = txType == IDTX ? TX_CLASS_2D : tx_class(txType);
    if (cls == TX_CLASS_VERT)
        return t == TX_4X4 ? Mrow_Scan_4x4 : t == TX_4X8 ? Mrow_Scan_4x8 : t == TX_8X4 ? Mrow_Scan_8x4 : t == TX_8X8 ? Mrow_Scan_8x8
             : t == TX_8X16 ? Mrow_Scan_8x16 : t == TX_16X8 ? Mrow_Scan_16x8 : t == TX_16X16 ? Mrow_Scan_16x16
             : t == TX_4X16 ? Mrow_Scan_4x16 : Mrow_Scan_16x4;
    if (cls == TX_CLASS_HORIZ)
        return t == TX_4X4 ? Mcol_Scan_4x4 : t == TX_4X8 ? Mcol_Scan_4x8 : t == TX_8X4 ? Mcol_Scan_8x4 : t == TX_8X8 ? Mcol_Scan_8x8
             : t == TX_8X16 ? Mcol_Scan_8x16 : t == TX_16X8 ? Mcol_Scan_16x8 : t == TX_16X16 ? Mcol_Scan_16x16
             : t == TX_4X16 ? Mcol_Scan_4x16 : Mcol_Scan_16x4;
    return t == TX_4X4 ? Default_Scan_4x4 : t == TX_4X8 ? Default_Scan_4x8 : t == TX_8X4 ? Default_Scan_8x4
         : t == TX_8X8 ? Default_Scan_8x8 : t == TX_8X16 ? Default_Scan_8x16 : t == TX_16X8 ? Default_Scan_16x8
         : t == TX_16X16 ? Default_Scan_16x16 : t == TX_16X32 ? Default_Scan_16x32 : t == TX_32X16 ? Default_Scan_32x16
         : t == TX_4X16 ? Default_Scan_4x16 : t == TX_16X4 ? Default_Scan_16x4 : t == TX_8X32 ? Default_Scan_8x32
         : t == TX_32X8 ? Default_Scan_32x8 : Default_Scan_32x32;
}

// the default (zig-zag) scans of the tx sizes without a 64 side, concatenated in tx-size order: offset of size t
constexpr int DSCAN_TOTAL = 3344;
AV1M_HD constexpr int dscan_off(int t) {
    return t == TX_4X4 ? 0 : t == TX_8X8 ? 16 : t == TX_16X16 ? 80 : t == TX_32X32 ? 336 : t == TX_4X8 ? 1360 : t == TX_8X4 ? 1392
         : t == TX_8X16 ? 1424 : t == TX_16X8 ? 1552 : t == TX_16X32 ? 1680 : t == TX_32X16 ? 2192 : t == TX_4X16 ? 2704
         : t == TX_16X4 ? 2768 : t == TX_8X32 ? 2832 : 3088;
}
// the coefficient context of tx size t (get_tx_size_context)
AV1M_HD constexpr int txsz_ctx_of(int t) { return (av1t::Tx_Size_Sqr[t] + av1t::Tx_Size_Sqr_Up[t] + 1) >> 1; }

// ---- block geometry ----
AV1M_HD int has_chroma(int bs, int mi_row, int mi_col, int mono, int ss) {
    if (mono) return 0;
    if (ss && bh4(bs) == 1 && (mi_row & 1) == 0) return 0;
    if (ss && bw4(bs) == 1 && (mi_col & 1) == 0) return 0;
    return 1;
}
AV1M_HD int uv_bsize(int bs, int ss) { return av1t::Subsampled_Size[bs][ss][ss]; }
AV1M_HD int uv_tx_size(int bs, int ss) {
    const int t = av1t::Max_Tx_Size_Rect[uv_bsize(bs, ss)];
    const int w = tx_w(t), h = tx_h(t);
    if (w == 64 || h == 64) return w == 16 ? TX_16X32 : h == 16 ? TX_32X16 : TX_32X32;
    return t;
}
AV1M_HD int cfl_allowed(int bs) { return bw4(bs) <= 8 && bh4(bs) <= 8; }

// ---- exact 1-D inverse transforms (spec 7.13.2), register resident: every index is a compile-time constant once the
// loops over N are unrolled.  err collects the conformance flags (1: butterfly range, 2: ADST4 range).
struct Inv {
    static AV1M_HD int cos128(int a) {
        a &= 255;
        return a <= 64 ? av1t::Cos128_Lookup[a] : a <= 128 ? -av1t::Cos128_Lookup[128 - a]
             : a <= 192 ? -av1t::Cos128_Lookup[a - 128] : av1t::Cos128_Lookup[256 - a];
    }
    // 32-bit arithmetic (r <= 18: 8- and 10-bit).  While every value a step reads lies inside its range, the products
    // (at most 2^17 x 2^14) and sums stay below 2^31 and the results are exact; the first value out of its range is
    // then also exact, so it raises the range flag whose pass is discarded (the coefficients are halved and the
    // transform runs again), and what the discarded pass computes after it does not matter.  mul / add wrap (unsigned
    // arithmetic) so that such a pass has defined behaviour.
    static AV1M_HD int32_t mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
    static AV1M_HD int32_t add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
    static AV1M_HD int32_t sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
    static AV1M_HD int32_t rnd12(int32_t v) { return add(v, 2048) >> 12; }
    static AV1M_HD void range(int32_t v, int bits, uint32_t &err, uint32_t f) {
        if (v < -(1 << (bits - 1)) || v > (1 << (bits - 1)) - 1) err |= f;
    }
    static AV1M_HD void B(int32_t *T, int a, int b, int ang, int flip, int r, uint32_t &err) {
        const int32_t c = cos128(ang), s = cos128(ang - 64);
        const int32_t x = rnd12(sub(mul(T[a], c), mul(T[b], s))), y = rnd12(add(mul(T[a], s), mul(T[b], c)));
        range(x, r, err, 1u);
        range(y, r, err, 1u);
        T[a] = flip ? y : x;
        T[b] = flip ? x : y;
    }
    static AV1M_HD void H(int32_t *T, int a, int b, int flip, int r) {
        const int aa = flip ? b : a, bb = flip ? a : b;
        const int32_t x = T[aa], y = T[bb], lo = -(1 << (r - 1)), hi = (1 << (r - 1)) - 1;
        const int32_t s = add(x, y), d = sub(x, y);
        T[aa] = s < lo ? lo : s > hi ? hi : s;
        T[bb] = d < lo ? lo : d > hi ? hi : d;
    }
    static constexpr int brev(int bits, int x) {
        int t = 0;
        for (int i = 0; i < bits; i++) t |= ((x >> i) & 1) << (bits - 1 - i);
        return t;
    }
    template <int n> static AV1M_HD void dct(int32_t *T, int r, uint32_t &err) {
        constexpr int n0 = 1 << n;
        int32_t c[n0];
#pragma unroll
        for (int i = 0; i < n0; i++) c[i] = T[i];
#pragma unroll
        for (int i = 0; i < n0; i++) T[i] = c[brev(n, i)];
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 16; i++) B(T, 32 + i, 63 - i, 63 - 4 * brev(4, i), 0, r, err);
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 8; i++) B(T, 16 + i, 31 - i, 6 + (brev(3, 7 - i) << 3), 0, r, err);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 16; i++) H(T, 32 + i * 2, 33 + i * 2, i & 1, r);
        }
        if (n >= 4) {
#pragma unroll
            for (int i = 0; i < 4; i++) B(T, 8 + i, 15 - i, 12 + (brev(2, 3 - i) << 4), 0, r, err);
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 8; i++) H(T, 16 + 2 * i, 17 + 2 * i, i & 1, r);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) B(T, 62 - i * 4 - j, 33 + i * 4 + j, 60 - 16 * brev(2, i) + 64 * j, 1, r, err);
        }
        if (n >= 3) {
#pragma unroll
            for (int i = 0; i < 2; i++) B(T, 4 + i, 7 - i, 56 - 32 * i, 0, r, err);
        }
        if (n >= 4) {
#pragma unroll
            for (int i = 0; i < 4; i++) H(T, 8 + 2 * i, 9 + 2 * i, i & 1, r);
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) B(T, 30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1, r, err);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) H(T, 32 + i * 4 + j, 35 + i * 4 - j, i & 1, r);
        }
#pragma unroll
        for (int i = 0; i < 2; i++) B(T, 2 * i, 2 * i + 1, 32 + 16 * i, 1 - i, r, err);
        if (n >= 3) {
#pragma unroll
            for (int i = 0; i < 2; i++) H(T, 4 + 2 * i, 5 + 2 * i, i, r);
        }
        if (n >= 4) {
#pragma unroll
            for (int i = 0; i < 2; i++) B(T, 14 - i, 9 + i, 48 + 64 * i, 1, r, err);
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) H(T, 16 + 4 * i + j, 19 + 4 * i - j, i & 1, r);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) B(T, 61 - i * 8 - j, 34 + i * 8 + j, 56 - i * 32 + (j >> 1) * 64, 1, r, err);
        }
#pragma unroll
        for (int i = 0; i < 2; i++) H(T, i, 3 - i, 0, r);
        if (n >= 3) B(T, 6, 5, 32, 1, r, err);
        if (n >= 4) {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) H(T, 8 + 4 * i + j, 11 + 4 * i - j, i, r);
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 4; i++) B(T, 29 - i, 18 + i, 48 + (i >> 1) * 64, 1, r, err);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) H(T, 32 + 8 * i + j, 39 + 8 * i - j, i & 1, r);
        }
        if (n >= 3) {
#pragma unroll
            for (int i = 0; i < 4; i++) H(T, i, 7 - i, 0, r);
        }
        if (n >= 4) {
#pragma unroll
            for (int i = 0; i < 2; i++) B(T, 13 - i, 10 + i, 32, 1, r, err);
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) H(T, 16 + i * 8 + j, 23 + i * 8 - j, i, r);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 8; i++) B(T, 59 - i, 36 + i, i < 4 ? 48 : 112, 1, r, err);
        }
        if (n >= 4) {
#pragma unroll
            for (int i = 0; i < 8; i++) H(T, i, 15 - i, 0, r);
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 4; i++) B(T, 27 - i, 20 + i, 32, 1, r, err);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 8; i++) { H(T, 32 + i, 47 - i, 0, r); H(T, 48 + i, 63 - i, 1, r); }
        }
        if (n >= 5) {
#pragma unroll
            for (int i = 0; i < 16; i++) H(T, i, 31 - i, 0, r);
        }
        if (n == 6) {
#pragma unroll
            for (int i = 0; i < 8; i++) B(T, 55 - i, 40 + i, 32, 1, r, err);
#pragma unroll
            for (int i = 0; i < 32; i++) H(T, i, 63 - i, 0, r);
        }
    }
    template <int n> static AV1M_HD void adst_in(int32_t *T) {
        constexpr int n0 = 1 << n;
        int32_t c[n0];
#pragma unroll
        for (int i = 0; i < n0; i++) c[i] = T[i];
#pragma unroll
        for (int i = 0; i < n0; i++) T[i] = c[(i & 1) ? i - 1 : n0 - i - 1];
    }
    template <int n> static AV1M_HD void adst_out(int32_t *T) {
        constexpr int n0 = 1 << n;
        int32_t c[n0];
#pragma unroll
        for (int i = 0; i < n0; i++) c[i] = T[i];
#pragma unroll
        for (int i = 0; i < n0; i++) {
            const int a = (i >> 3) & 1, b = ((i >> 2) & 1) ^ ((i >> 3) & 1), cc = ((i >> 1) & 1) ^ ((i >> 2) & 1), d = (i & 1) ^ ((i >> 1) & 1);
            const int idx = ((d << 3) | (cc << 2) | (b << 1) | a) >> (4 - n);
            T[i] = (i & 1) ? -c[idx] : c[idx];
        }
    }
    static AV1M_HD void adst4(int32_t *T, int r, uint32_t &err) {
        constexpr int32_t S1 = 1321, S2 = 2482, S3 = 3344, S4 = 3803;
        int32_t s0 = mul(S1, T[0]), s1 = mul(S2, T[0]), s2 = mul(S3, T[1]), s3 = mul(S4, T[2]), s4 = mul(S1, T[2]), s5 = mul(S2, T[3]),
                s6 = mul(S4, T[3]);
        const int32_t a7 = sub(T[0], T[2]), b7 = add(a7, T[3]);
        range(a7, r + 1, err, 2u); range(b7, r, err, 2u);
        range(s0, r + 12, err, 2u); range(s1, r + 12, err, 2u); range(s2, r + 12, err, 2u); range(s3, r + 12, err, 2u);
        range(s4, r + 12, err, 2u); range(s5, r + 12, err, 2u); range(s6, r + 12, err, 2u);
        s0 = add(s0, s3); s1 = sub(s1, s4); s3 = s2; s2 = mul(S3, b7);
        range(s0, r + 12, err, 2u); range(s1, r + 12, err, 2u); range(s2, r + 12, err, 2u);
        s0 = add(s0, s5); s1 = sub(s1, s6);
        range(s0, r + 12, err, 2u); range(s1, r + 12, err, 2u);
        int32_t x0 = add(s0, s3), x1 = add(s1, s3), x2 = s2, x3 = add(s0, s1);
        range(x3, r + 12, err, 2u);
        x3 = sub(x3, s3);
        range(x0, r + 12, err, 2u); range(x1, r + 12, err, 2u); range(x2, r + 12, err, 2u); range(x3, r + 12, err, 2u);
        T[0] = rnd12(x0); T[1] = rnd12(x1); T[2] = rnd12(x2); T[3] = rnd12(x3);
    }
    static AV1M_HD void adst8(int32_t *T, int r, uint32_t &err) {
        adst_in<3>(T);
#pragma unroll
        for (int i = 0; i < 4; i++) B(T, 2 * i, 2 * i + 1, 60 - 16 * i, 1, r, err);
#pragma unroll
        for (int i = 0; i < 4; i++) H(T, i, 4 + i, 0, r);
#pragma unroll
        for (int i = 0; i < 2; i++) B(T, 4 + 3 * i, 5 + i, 48 - 32 * i, 1, r, err);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) H(T, 4 * j + i, 2 + 4 * j + i, 0, r);
#pragma unroll
        for (int i = 0; i < 2; i++) B(T, 2 + 4 * i, 3 + 4 * i, 32, 1, r, err);
        adst_out<3>(T);
    }
    static AV1M_HD void adst16(int32_t *T, int r, uint32_t &err) {
        adst_in<4>(T);
#pragma unroll
        for (int i = 0; i < 8; i++) B(T, 2 * i, 2 * i + 1, 62 - 8 * i, 1, r, err);
#pragma unroll
        for (int i = 0; i < 8; i++) H(T, i, 8 + i, 0, r);
#pragma unroll
        for (int i = 0; i < 2; i++) { B(T, 8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1, r, err); B(T, 13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1, r, err); }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) H(T, 8 * j + i, 4 + 8 * j + i, 0, r);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) B(T, 4 + 8 * j + 3 * i, 5 + 8 * j + i, 48 - 32 * i, 1, r, err);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) H(T, 4 * j + i, 2 + 4 * j + i, 0, r);
#pragma unroll
        for (int i = 0; i < 4; i++) B(T, 2 + 4 * i, 3 + 4 * i, 32, 1, r, err);
        adst_out<4>(T);
    }
    template <int n> static AV1M_HD void idt(int32_t *T) {
#pragma unroll
        for (int i = 0; i < (1 << n); i++) {
            if (n == 2) T[i] = rnd12(mul(T[i], 5793));
            else if (n == 3) T[i] = mul(T[i], 2);
            else if (n == 4) T[i] = rnd12(mul(T[i], 11586));
            else T[i] = mul(T[i], 4);
        }
    }
    template <int n> static AV1M_HD void run(int32_t *T, int kind, int r, uint32_t &err) {
        if (kind == K_DCT) dct<n>(T, r, err);
        else if (kind == K_ADST) {
            if constexpr (n == 2) adst4(T, r, err);
            else if constexpr (n == 3) adst8(T, r, err);
            else if constexpr (n == 4) adst16(T, r, err);
        } else {
            if constexpr (n <= 5) idt<n>(T);
        }
    }
};

// runtime log2 dispatch of one 1-D inverse transform
AV1M_HD void inv_1d(int32_t *T, int kind, int n, int r, uint32_t &err) {
    switch (n) {
    case 2: Inv::run<2>(T, kind, r, err); break;
    case 3: Inv::run<3>(T, kind, r, err); break;
    case 4: Inv::run<4>(T, kind, r, err); break;
    case 5: Inv::run<5>(T, kind, r, err); break;
    default: Inv::run<6>(T, kind, r, err); break;
    }
}

// ---- intra prediction (spec 7.11.2) ----
AV1M_HD const uint8_t *sm_weights(int l) {
    return l == 2 ? av1t::Sm_Weights_Tx_4x4 : l == 3 ? av1t::Sm_Weights_Tx_8x8 : l == 4 ? av1t::Sm_Weights_Tx_16x16
         : l == 5 ? av1t::Sm_Weights_Tx_32x32 : av1t::Sm_Weights_Tx_64x64;
}
// one predicted sample (row i, column j) of a non-DC mode from AboveRow / LeftCol (index 0 = first sample, -1 the
// corner; after upsampling index -2 exists)
template <typename P>
AV1M_HD int pred_px(int mode, int pa, int i, int j, int lw, int lh, const P *A, const P *L, int upA, int upL) {
    const int w = 1 << lw, h = 1 << lh;
    if (is_dir(mode)) {
        if (pa == 90) return A[j];
        if (pa == 180) return L[i];
        if (pa < 90) {
            const int dx = av1t::Dr_Intra_Derivative[pa], idx = (i + 1) * dx;
            const int base = (idx >> (6 - upA)) + (j << upA), sh = ((idx << upA) >> 1) & 0x1F, mb = (w + h - 1) << upA;
            return base < mb ? (A[base] * (32 - sh) + A[base + 1] * sh + 16) >> 5 : A[mb];
        }
        if (pa < 180) {
            const int dx = av1t::Dr_Intra_Derivative[180 - pa];
            int idx = (j << 6) - (i + 1) * dx;
            if ((idx >> (6 - upA)) >= -(1 << upA)) {
                const int base = idx >> (6 - upA), sh = ((idx * (1 << upA)) >> 1) & 0x1F;
                return (A[base] * (32 - sh) + A[base + 1] * sh + 16) >> 5;
            }
            const int dy = av1t::Dr_Intra_Derivative[pa - 90];
            idx = (i << 6) - (j + 1) * dy;
            const int base = idx >> (6 - upL), sh = ((idx * (1 << upL)) >> 1) & 0x1F;
            return (L[base] * (32 - sh) + L[base + 1] * sh + 16) >> 5;
        }
        const int dy = av1t::Dr_Intra_Derivative[270 - pa], idx = (j + 1) * dy;
        const int base = (idx >> (6 - upL)) + (i << upL), sh = ((idx << upL) >> 1) & 0x1F;
        return (L[base] * (32 - sh) + L[base + 1] * sh + 16) >> 5;
    }
    if (mode == SMOOTH_PRED) {
        const uint8_t *wx = sm_weights(lw), *wy = sm_weights(lh);
        return (wy[i] * A[j] + (256 - wy[i]) * L[h - 1] + wx[j] * L[i] + (256 - wx[j]) * A[w - 1] + 256) >> 9;
    }
    if (mode == SMOOTH_V_PRED) {
        const uint8_t *wy = sm_weights(lh);
        return (wy[i] * A[j] + (256 - wy[i]) * L[h - 1] + 128) >> 8;
    }
    if (mode == SMOOTH_H_PRED) {
        const uint8_t *wx = sm_weights(lw);
        return (wx[j] * L[i] + (256 - wx[j]) * A[w - 1] + 128) >> 8;
    }
    const int a = A[j], l = L[i], tl = A[-1], base = a + l - tl;   // PAETH
    const int pL = base > l ? base - l : l - base, pT = base > a ? base - a : a - base, pTL = base > tl ? base - tl : tl - base;
    return (pL <= pT && pL <= pTL) ? l : pT <= pTL ? a : tl;
}
// CH predicted samples of column j, rows i0, i0 + di, ..., i0 + (CH - 1) di, of any mode (DC_PRED: dcv): pred_px's
// cases with the mode and zone tests made once for the CH samples, so a group whose lanes share the mode branches once
// per run of samples instead of once per sample
template <int CH, typename P>
AV1M_HD void pred_col(int mode, int pa, int dcv, int i0, int di, int j, int lw, int lh, const P *A, const P *L, int upA, int upL,
                      int (&o)[CH]) {
    const int w = 1 << lw, h = 1 << lh;
    if (mode == DC_PRED) {
        for (int c = 0; c < CH; c++) o[c] = dcv;
    } else if (is_dir(mode)) {
        if (pa == 90) {
            const int v = A[j];
            for (int c = 0; c < CH; c++) o[c] = v;
        } else if (pa == 180) {
            for (int c = 0; c < CH; c++) o[c] = L[i0 + c * di];
        } else if (pa < 90) {
            const int dx = av1t::Dr_Intra_Derivative[pa], mb = (w + h - 1) << upA, am = A[mb];
            for (int c = 0; c < CH; c++) {
                const int idx = (i0 + c * di + 1) * dx;
                const int base = (idx >> (6 - upA)) + (j << upA), sh = ((idx << upA) >> 1) & 0x1F;
                o[c] = base < mb ? (A[base] * (32 - sh) + A[base + 1] * sh + 16) >> 5 : am;
            }
        } else if (pa < 180) {
            const int dx = av1t::Dr_Intra_Derivative[180 - pa], dy = av1t::Dr_Intra_Derivative[pa - 90];
            for (int c = 0; c < CH; c++) {
                const int i = i0 + c * di;
                const int ia = (j << 6) - (i + 1) * dx;
                if ((ia >> (6 - upA)) >= -(1 << upA)) {
                    const int base = ia >> (6 - upA), sh = ((ia * (1 << upA)) >> 1) & 0x1F;
                    o[c] = (A[base] * (32 - sh) + A[base + 1] * sh + 16) >> 5;
                } else {
                    const int il = (i << 6) - (j + 1) * dy;
                    const int base = il >> (6 - upL), sh = ((il * (1 << upL)) >> 1) & 0x1F;
                    o[c] = (L[base] * (32 - sh) + L[base + 1] * sh + 16) >> 5;
                }
            }
        } else {
            const int dy = av1t::Dr_Intra_Derivative[270 - pa], idx = (j + 1) * dy, sh = ((idx << upL) >> 1) & 0x1F;
            for (int c = 0; c < CH; c++) {
                const int base = (idx >> (6 - upL)) + ((i0 + c * di) << upL);
                o[c] = (L[base] * (32 - sh) + L[base + 1] * sh + 16) >> 5;
            }
        }
    } else if (mode == PAETH_PRED) {
        const int a = A[j], tl = A[-1];
        for (int c = 0; c < CH; c++) {
            const int l = L[i0 + c * di], base = a + l - tl;
            const int pL = base > l ? base - l : l - base, pT = base > a ? base - a : a - base, pTL = base > tl ? base - tl : tl - base;
            o[c] = (pL <= pT && pL <= pTL) ? l : pT <= pTL ? a : tl;
        }
    } else {
        // the SMOOTH modes: (x + 128) >> 8 = (2x + 256) >> 9, so SMOOTH_V / SMOOTH_H are SMOOTH with the other
        // direction's pair replaced by a second copy of their own
        const uint8_t *wx = sm_weights(lw), *wy = sm_weights(lh);
        const int a = A[j], lb = L[h - 1], ar = A[w - 1], wj = wx[j];
        for (int c = 0; c < CH; c++) {
            const int i = i0 + c * di, wi = wy[i], l = L[i];
            const int vv = wi * a + (256 - wi) * lb, hh = wj * l + (256 - wj) * ar;
            o[c] = ((mode == SMOOTH_PRED ? vv + hh : mode == SMOOTH_V_PRED ? 2 * vv : 2 * hh) + 256) >> 9;
        }
    }
}
AV1M_HD int edge_strength(int w, int h, int ft, int delta) {
    const int d = delta < 0 ? -delta : delta, s = w + h;
    if (ft == 0) {
        if (s <= 8) return d >= 56 ? 1 : 0;
        if (s <= 16) return d >= 40 ? 1 : 0;
        if (s <= 24) return d >= 32 ? 3 : d >= 16 ? 2 : d >= 8 ? 1 : 0;
        if (s <= 32) return d >= 32 ? 3 : d >= 4 ? 2 : 1;
        return 3;
    }
    if (s <= 8) return d >= 64 ? 2 : d >= 40 ? 1 : 0;
    if (s <= 16) return d >= 48 ? 2 : d >= 20 ? 1 : 0;
    if (s <= 24) return d >= 4 ? 3 : 0;
    return 3;
}
AV1M_HD int edge_upsample(int w, int h, int ft, int delta) {
    const int d = delta < 0 ? -delta : delta;
    if (d <= 0 || d >= 40) return 0;
    return ft == 0 ? (w + h <= 16) : (w + h <= 8);
}

// ---- structures shared by the host and the kernels ----
// symbol costs in bits of one coefficient-CDF q-context (-log2 of the default CDF probabilities)
struct CostTab {
    float y_mode[5][5][13];
    float uv_cfl_no[13][13];
    float uv_cfl_yes[13][14];
    float angle_delta[8][7];
    float skip[3][2];
    float tx_set1[2][13][7];
    float tx_set2[3][13][5];
    float cfl_sign[8];
    float cfl_alpha[6][16];
    float txb_skip[5][13][2];
    float eob_pt16[2][2][5], eob_pt32[2][2][6], eob_pt64[2][2][7], eob_pt128[2][2][8], eob_pt256[2][2][9];
    float eob_pt512[2][10], eob_pt1024[2][11];
    float eob_extra[5][2][9][2];
    float dc_sign[2][3][2];
    float coeff_base_eob[5][2][4][3];
    float coeff_base[5][2][42][4];
    float coeff_br[5][2][21][4];
    float tx8[3][2], tx16[3][3], tx32[3][3], tx64[3][3];
};
// device-resident tables, built once on the host (tables.cpp) and uploaded
struct DevTables {
    float fwdT[FWD_FLOATS];      // forward matrices, transposed: fwdT[off + k * n + i] = F[i][k]
    float fwd[FWD_FLOATS];       // row-major F[i][k] (identity kinds read the diagonal)
    int fwdOff[3][7];
    float txg[19][16];
    float txScale[19];
    float avgYMode[13];
    float avgPart[5][10];
    float avgSkip0;
    float lambda[2][256];        // [bd == 10][qindex], lambdaMul 1
    float lamS[2][256];          // sqrt(lambda): the screening weight with satdLambda 1
    CostTab cost[4];
    uint16_t iscan[DSCAN_TOTAL]; // inverse default scans: iscan[dscan_off(t) + position] = scan index
};

// tunables of the search (the reference's GpParams, gp.h:29-68, for the knobs this core implements)
struct SearchParams {
    float lambdaMul, satdLambda, rnd, zbin, rndC, zbinC;
    int K, Krecheck, txSearch, txSearchCL, txAllCL, deltas, uvDeltas, cfl, uvK, shapeMask, partMask8, partMaskN, maxPart;
    float uvOverheadBits, splitBias, lambdaCMul;
};

// per-image descriptor in device memory
struct Img {
    int W, H, MiCols, MiRows, sbCols, sbRows, aw, ah, ss, mono, bd, qindex, qctx, edgeFilter;
    int tileCols, tileRows;
    int tileColStart[65], tileRowStart[65];   // mi units
    uint32_t shOff[NSHAPES];                  // first analysis record of each shape
    int shNbx[NSHAPES], shNby[NSHAPES];
    uint32_t sbBase;                          // first SB of the image in the batch's SB arrays
    const void *src[3];                       // padded source planes (aw x ah; chroma aw >> ss x ah >> ss), uint8 / uint16
    int spitch[3];                            // source pitches, samples
    void *rec[3];                             // reconstruction planes, same geometry
    int pitch[3];                             // reconstruction pitches, samples
    uint8_t *mY, *mUV, *mSkip, *mMiSize, *mTx; // 4x4 maps (MiRows x MiCols)
    uint8_t *cLvl[3], *cDc[3];                // coefficient-context maps per plane (plane units)
    // post-filters: effective loop filter levels (Y vertical, Y horizontal, U, V; 0 off) and sharpness; CDEF with one
    // strength set (secondary strengths 0, 1, 2, 4), the deblocked frame filtered into cdefOut (same geometry as rec)
    uint8_t lfLevel[4], lfSharp, cdefOn, cdefDamping, cdefYPri, cdefYSec, cdefUVPri, cdefUVSec;
    void *cdefOut[3];
};
struct BlockRes { float J; uint8_t cand[MAXK]; uint8_t ncand; uint8_t tx; };
struct Leaf { uint8_t bsize, x4, y4, pad; };
struct SbOut { uint32_t nBlocks, nTxb, nCoeffs, pad; };
// av1g_block / av1g_txb (the DecisionRecord's 16-byte entries)
struct RecBlock { uint16_t mi_row, mi_col; uint8_t bsize, skip, tx_depth, y_mode; int8_t angle_delta_y; uint8_t uv_mode;
                  int8_t angle_delta_uv, cfl_alpha_u, cfl_alpha_v; uint8_t filter_intra; int8_t cdef_idx; uint8_t sb_qindex; };
struct RecTxb { uint32_t coeff_off; uint16_t x4, y4; uint8_t plane, tx_size, tx_type, flags; uint16_t eob, reserved; };
struct PackOff { uint32_t gBlk, gTxb, rTxb, rCoef; uint64_t gCoef; };

// per-batch device pointers
struct Batch {
    int nimg;
    const Img *img;
    const DevTables *T;
    BlockRes *blk;
    const uint32_t *shStart;     // [NSHAPES][nimg + 1] first record of each image in each shape
    uint32_t nSb;
    const uint32_t *sbImg;       // [nSb]
    uint16_t *leafCount;         // [nSb]
    Leaf *leaves;                // [nSb * MAX_LEAVES]
    float *sbJ;                  // [nSb]
    SbOut *sbOut;                // [nSb]
    RecBlock *oBlocks;           // [nSb * MAX_LEAVES]
    uint16_t *oBlockTxb;         // [nSb * MAX_LEAVES]
    RecTxb *oTxb;                // [nSb * txbCap]
    int16_t *oCoeffs;            // [nSb * coefCap]
    uint32_t txbCap, coefCap;    // per-SB record slots: SB_TXB_CAP / SB_COEF_CAP, or the 4:4:4 ones
    float *dpRel;                // [nSb][255] or null: per node, the DP's relative cost increase without its children
                                 // (0..84), without its own shapes (85..169), without its 4-way shapes (170..254)
    uint32_t *stats;             // [8]: blocks, skip blocks, CfL blocks, range repairs
    const uint32_t *lnMask;      // [nSb][12] learned pruning: expanded nodes (words 0-2), screened luma modes (3), nodes
                                 // keeping their own shapes (4-6), their 4-way shapes (8-10); null = exhaustive
    unsigned long long *flops;   // [64] or null: forward-transform FLOPs (fwd_flops) of the evaluations, analysis in
                                 // [0, 32), closed loop in [32, 64) (spread over slots to keep the atomics apart)
};

// an external screener's candidate (GpExtCand, encoder/src/gp.h:136-143; av1mi_ext_cand)
struct ExtCand { float J; uint8_t mode; int8_t delta; uint8_t tx, pad; };
// the analysis hooks of one shape: the external candidates of its blocks (null: screening) and the list of blocks to
// analyse (null: all); GpBatch::extCand / extOff / extTopK (gp.h:164-171), GpAnX (an_stage.h:23-27)
struct AnHook {
    const ExtCand *ext;          // [blocks][extTopK], entry of block (img, local) at (extOff[img] + local) * extTopK
    const int64_t *extOff;       // [nimg] this shape's first entry per image
    int extTopK;
    const uint32_t *list;        // shape-global block indices (shStart order)
    const uint32_t *count;       // [1] entries of list
};
// host side: the hooks of every shape (null: none) and each shape's first record in Batch::blk
struct AnHooks {
    const AnHook *hook;
    const uint32_t *base;
};

// the device work queue of the persistent closed-loop schedule
struct RcQueue {
    const uint32_t *order;       // [n] superblocks in anti-diagonal wavefront order (step, image)
    uint32_t n;
    uint32_t *head;              // next item to take
    uint32_t *done;              // [nSb] set when a superblock is complete
    uint32_t *abort;             // raised by a wait that ran out of time
    uint64_t spinTicks;          // wait limit in wall-clock ticks
};

}  // namespace av1m
