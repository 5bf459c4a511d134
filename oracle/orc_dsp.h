// Oracle (test infrastructure only): CPU restatement of the AV1 spec integer arithmetic that the intra path uses,
// written serially, one function per spec process, following the reference's av1core/include/av1g/dsp.h and
// blockinfo.h.  Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it, as the checker.
// Every function names the reference file:line it restates.
#pragma once
#include <stdint.h>
#include <string.h>
#include "../av1gpu_amd/csrc/av1_spec_tables.h"

namespace orc {

// ---- symbol values (AV1 spec section 6.10; reference enums_gen.h) ----
enum { BLOCK_4X4 = 0, BLOCK_4X8, BLOCK_8X4, BLOCK_8X8, BLOCK_8X16, BLOCK_16X8, BLOCK_16X16, BLOCK_16X32, BLOCK_32X16,
       BLOCK_32X32, BLOCK_32X64, BLOCK_64X32, BLOCK_64X64, BLOCK_4X16 = 16, BLOCK_16X4, BLOCK_8X32, BLOCK_32X8,
       BLOCK_16X64, BLOCK_64X16, BLOCK_INVALID };
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
enum { ERR_B_RANGE = 1u, ERR_ADST4_RANGE = 2u };
static const int EDGE_OFF = 16, EDGE_LEN = 16 + 2 * 64 + 16;   // dsp.h:521-522

inline int clip3(int lo, int hi, int x) { return x < lo ? lo : x > hi ? hi : x; }
inline int clip1(int x, int bd) { return clip3(0, (1 << bd) - 1, x); }
inline int32_t round2(int32_t x, int n) { return n == 0 ? x : (x + (1 << (n - 1))) >> n; }
inline int64_t round2_64(int64_t x, int n) { return n == 0 ? x : (x + ((int64_t)1 << (n - 1))) >> n; }
inline int32_t round2signed(int32_t x, int n) { return x >= 0 ? round2(x, n) : -round2(-x, n); }

inline int bw4(int bs) { return av1t::Num_4x4_Blocks_Wide[bs]; }
inline int bh4(int bs) { return av1t::Num_4x4_Blocks_High[bs]; }
inline int txw(int t) { return av1t::Tx_Width[t]; }
inline int txh(int t) { return av1t::Tx_Height[t]; }
inline int tx_coeffs(int t) { return (txw(t) > 32 ? 32 : txw(t)) * (txh(t) > 32 ? 32 : txh(t)); }

// get_tx_set for intra blocks (dsp.h:55-62)
inline int tx_set_intra(int txSz) {
    const int sqr = av1t::Tx_Size_Sqr[txSz], up = av1t::Tx_Size_Sqr_Up[txSz];
    if (up >= TX_32X32) return TX_SET_DCTONLY;
    if (sqr == TX_16X16) return TX_SET_INTRA_2;
    return TX_SET_INTRA_1;
}
inline int tx_in_set(int set, int t) { return av1t::Tx_Type_In_Set_Intra[set][t]; }
// compute_tx_type for an intra chroma tx block (dsp.h:67-72)
inline int chroma_tx_type(int uvTx, int uvMode) {
    if (av1t::Tx_Size_Sqr_Up[uvTx] > TX_32X32) return DCT_DCT;
    const int t = av1t::Mode_To_Txfm[uvMode];
    return tx_in_set(tx_set_intra(uvTx), t) ? t : DCT_DCT;
}
// dsp.h:75-79
inline int tx_class(int t) {
    if (t == V_DCT || t == V_ADST || t == V_FLIPADST) return TX_CLASS_VERT;
    if (t == H_DCT || t == H_ADST || t == H_FLIPADST) return TX_CLASS_HORIZ;
    return TX_CLASS_2D;
}
// dsp.h:369-388
inline int row_kind(int t) {
    switch (t) {
    case DCT_DCT: case ADST_DCT: case FLIPADST_DCT: case H_DCT: return K_DCT;
    case IDTX: case V_DCT: case V_ADST: case V_FLIPADST: return K_IDT;
    default: return K_ADST;
    }
}
inline int col_kind(int t) {
    switch (t) {
    case DCT_DCT: case DCT_ADST: case DCT_FLIPADST: case V_DCT: return K_DCT;
    case IDTX: case H_DCT: case H_ADST: case H_FLIPADST: return K_IDT;
    default: return K_ADST;
    }
}
inline int flip_ud(int t) { return t == FLIPADST_DCT || t == FLIPADST_ADST || t == V_FLIPADST || t == FLIPADST_FLIPADST; }
inline int flip_lr(int t) { return t == DCT_FLIPADST || t == ADST_FLIPADST || t == H_FLIPADST || t == FLIPADST_FLIPADST; }

// get_scan (dsp.h:82-131)
inline const uint16_t *get_scan(int txSz, int txType) {
    using namespace av1t;
    if (txSz == TX_16X64) return Default_Scan_16x32;
    if (txSz == TX_64X16) return Default_Scan_32x16;
    if (Tx_Size_Sqr_Up[txSz] == TX_64X64) return Default_Scan_32x32;
    const int cls = txType == IDTX ? TX_CLASS_2D : tx_class(txType);
    if (cls == TX_CLASS_VERT) {
        switch (txSz) {
        case TX_4X4: return Mrow_Scan_4x4; case TX_4X8: return Mrow_Scan_4x8; case TX_8X4: return Mrow_Scan_8x4;
        case TX_8X8: return Mrow_Scan_8x8; case TX_8X16: return Mrow_Scan_8x16; case TX_16X8: return Mrow_Scan_16x8;
        case TX_16X16: return Mrow_Scan_16x16; case TX_4X16: return Mrow_Scan_4x16; default: return Mrow_Scan_16x4;
        }
    }
    if (cls == TX_CLASS_HORIZ) {
        switch (txSz) {
        case TX_4X4: return Mcol_Scan_4x4; case TX_4X8: return Mcol_Scan_4x8; case TX_8X4: return Mcol_Scan_8x4;
        case TX_8X8: return Mcol_Scan_8x8; case TX_8X16: return Mcol_Scan_8x16; case TX_16X8: return Mcol_Scan_16x8;
        case TX_16X16: return Mcol_Scan_16x16; case TX_4X16: return Mcol_Scan_4x16; default: return Mcol_Scan_16x4;
        }
    }
    switch (txSz) {
    case TX_4X4: return Default_Scan_4x4; case TX_4X8: return Default_Scan_4x8; case TX_8X4: return Default_Scan_8x4;
    case TX_8X8: return Default_Scan_8x8; case TX_8X16: return Default_Scan_8x16; case TX_16X8: return Default_Scan_16x8;
    case TX_16X16: return Default_Scan_16x16; case TX_16X32: return Default_Scan_16x32; case TX_32X16: return Default_Scan_32x16;
    case TX_4X16: return Default_Scan_4x16; case TX_16X4: return Default_Scan_16x4; case TX_8X32: return Default_Scan_8x32;
    case TX_32X8: return Default_Scan_32x8; default: return Default_Scan_32x32;
    }
}

// ---- quantizer (spec 7.12.2; dsp.h:138-186) ----
inline int dc_q(int b, int bd) { return av1t::Dc_Qlookup[(bd - 8) >> 1][clip3(0, 255, b)]; }
inline int ac_q(int b, int bd) { return av1t::Ac_Qlookup[(bd - 8) >> 1][clip3(0, 255, b)]; }
inline int dq_denom(int txSz) {
    if (txSz == TX_32X32 || txSz == TX_16X32 || txSz == TX_32X16 || txSz == TX_16X64 || txSz == TX_64X16) return 2;
    if (txSz == TX_64X64 || txSz == TX_32X64 || txSz == TX_64X32) return 4;
    return 1;
}
// spec 7.12.3 step 1 (no quantizer matrix), dsp.h:151-158
inline int32_t dequant_coef(int32_t quant, int q, int den, int bd) {
    const int32_t dq = quant * q;
    int32_t a = (dq < 0 ? -dq : dq) & 0xFFFFFF;
    a /= den;
    if (dq < 0) a = -a;
    return clip3(-(1 << (7 + bd)), (1 << (7 + bd)) - 1, a);
}

// ---- 1-D inverse transforms (spec 7.13.2; dsp.h:216-413) ----
inline int brev(int bits, int x) {
    int t = 0;
    for (int i = 0; i < bits; i++) t |= ((x >> i) & 1) << (bits - 1 - i);
    return t;
}
inline int cos128(int angle) {
    const int a = angle & 255;
    if (a <= 64) return av1t::Cos128_Lookup[a];
    if (a <= 128) return -av1t::Cos128_Lookup[128 - a];
    if (a <= 192) return -av1t::Cos128_Lookup[a - 128];
    return av1t::Cos128_Lookup[256 - a];
}
inline int sin128(int angle) { return cos128(angle - 64); }
inline void chk(int64_t v, int bits, uint32_t *err, uint32_t f) {
    if (v < -((int64_t)1 << (bits - 1)) || v > ((int64_t)1 << (bits - 1)) - 1) *err |= f;
}
// butterfly rotation B(a, b, angle, flip) (spec 7.13.2.1)
inline void bfly(int32_t *T, int a, int b, int angle, int flip, int r, uint32_t *err) {
    const int64_t c = cos128(angle), s = sin128(angle);
    const int64_t x = (int64_t)T[a] * c - (int64_t)T[b] * s;
    const int64_t y = (int64_t)T[a] * s + (int64_t)T[b] * c;
    const int64_t xa = round2_64(x, 12), yb = round2_64(y, 12);
    chk(xa, r, err, ERR_B_RANGE);
    chk(yb, r, err, ERR_B_RANGE);
    if (flip) { T[a] = (int32_t)yb; T[b] = (int32_t)xa; } else { T[a] = (int32_t)xa; T[b] = (int32_t)yb; }
}
// Hadamard rotation H(a, b, flip) with saturation to r bits (spec 7.13.2.1)
inline void hbfly(int32_t *T, int a, int b, int flip, int r) {
    if (flip) { const int t = a; a = b; b = t; }
    const int64_t x = T[a], y = T[b], lo = -((int64_t)1 << (r - 1)), hi = ((int64_t)1 << (r - 1)) - 1;
    const int64_t s = x + y, d = x - y;
    T[a] = (int32_t)(s < lo ? lo : s > hi ? hi : s);
    T[b] = (int32_t)(d < lo ? lo : d > hi ? hi : d);
}
// inverse DCT of 2^n points (spec 7.13.2.3, steps 1-31)
inline void idct(int32_t *T, int n, int r, uint32_t *err) {
    int32_t c[64];
    const int n0 = 1 << n;
    for (int i = 0; i < n0; i++) c[i] = T[i];
    for (int i = 0; i < n0; i++) T[i] = c[brev(n, i)];
    if (n == 6) for (int i = 0; i < 16; i++) bfly(T, 32 + i, 63 - i, 63 - 4 * brev(4, i), 0, r, err);
    if (n >= 5) for (int i = 0; i < 8; i++) bfly(T, 16 + i, 31 - i, 6 + (brev(3, 7 - i) << 3), 0, r, err);
    if (n == 6) for (int i = 0; i < 16; i++) hbfly(T, 32 + i * 2, 33 + i * 2, i & 1, r);
    if (n >= 4) for (int i = 0; i < 4; i++) bfly(T, 8 + i, 15 - i, 12 + (brev(2, 3 - i) << 4), 0, r, err);
    if (n >= 5) for (int i = 0; i < 8; i++) hbfly(T, 16 + 2 * i, 17 + 2 * i, i & 1, r);
    if (n == 6) for (int i = 0; i < 4; i++) for (int j = 0; j < 2; j++)
        bfly(T, 62 - i * 4 - j, 33 + i * 4 + j, 60 - 16 * brev(2, i) + 64 * j, 1, r, err);
    if (n >= 3) for (int i = 0; i < 2; i++) bfly(T, 4 + i, 7 - i, 56 - 32 * i, 0, r, err);
    if (n >= 4) for (int i = 0; i < 4; i++) hbfly(T, 8 + 2 * i, 9 + 2 * i, i & 1, r);
    if (n >= 5) for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++)
        bfly(T, 30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1, r, err);
    if (n == 6) for (int i = 0; i < 8; i++) for (int j = 0; j < 2; j++) hbfly(T, 32 + i * 4 + j, 35 + i * 4 - j, i & 1, r);
    for (int i = 0; i < 2; i++) bfly(T, 2 * i, 2 * i + 1, 32 + 16 * i, 1 - i, r, err);
    if (n >= 3) for (int i = 0; i < 2; i++) hbfly(T, 4 + 2 * i, 5 + 2 * i, i, r);
    if (n >= 4) for (int i = 0; i < 2; i++) bfly(T, 14 - i, 9 + i, 48 + 64 * i, 1, r, err);
    if (n >= 5) for (int i = 0; i < 4; i++) for (int j = 0; j < 2; j++) hbfly(T, 16 + 4 * i + j, 19 + 4 * i - j, i & 1, r);
    if (n == 6) for (int i = 0; i < 2; i++) for (int j = 0; j < 4; j++)
        bfly(T, 61 - i * 8 - j, 34 + i * 8 + j, 56 - i * 32 + (j >> 1) * 64, 1, r, err);
    for (int i = 0; i < 2; i++) hbfly(T, i, 3 - i, 0, r);
    if (n >= 3) bfly(T, 6, 5, 32, 1, r, err);
    if (n >= 4) for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) hbfly(T, 8 + 4 * i + j, 11 + 4 * i - j, i, r);
    if (n >= 5) for (int i = 0; i < 4; i++) bfly(T, 29 - i, 18 + i, 48 + (i >> 1) * 64, 1, r, err);
    if (n == 6) for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) hbfly(T, 32 + 8 * i + j, 39 + 8 * i - j, i & 1, r);
    if (n >= 3) for (int i = 0; i < 4; i++) hbfly(T, i, 7 - i, 0, r);
    if (n >= 4) for (int i = 0; i < 2; i++) bfly(T, 13 - i, 10 + i, 32, 1, r, err);
    if (n >= 5) for (int i = 0; i < 2; i++) for (int j = 0; j < 4; j++) hbfly(T, 16 + i * 8 + j, 23 + i * 8 - j, i, r);
    if (n == 6) for (int i = 0; i < 8; i++) bfly(T, 59 - i, 36 + i, i < 4 ? 48 : 112, 1, r, err);
    if (n >= 4) for (int i = 0; i < 8; i++) hbfly(T, i, 15 - i, 0, r);
    if (n >= 5) for (int i = 0; i < 4; i++) bfly(T, 27 - i, 20 + i, 32, 1, r, err);
    if (n == 6) for (int i = 0; i < 8; i++) { hbfly(T, 32 + i, 47 - i, 0, r); hbfly(T, 48 + i, 63 - i, 1, r); }
    if (n >= 5) for (int i = 0; i < 16; i++) hbfly(T, i, 31 - i, 0, r);
    if (n == 6) for (int i = 0; i < 8; i++) bfly(T, 55 - i, 40 + i, 32, 1, r, err);
    if (n == 6) for (int i = 0; i < 32; i++) hbfly(T, i, 63 - i, 0, r);
}
// ADST input / output permutations (spec 7.13.2.4 / 7.13.2.5)
inline void adst_in_perm(int32_t *T, int n) {
    int32_t c[16];
    const int n0 = 1 << n;
    for (int i = 0; i < n0; i++) c[i] = T[i];
    for (int i = 0; i < n0; i++) T[i] = c[(i & 1) ? i - 1 : n0 - i - 1];
}
inline void adst_out_perm(int32_t *T, int n) {
    int32_t c[16];
    const int n0 = 1 << n;
    for (int i = 0; i < n0; i++) c[i] = T[i];
    for (int i = 0; i < n0; i++) {
        const int a = (i >> 3) & 1, b = ((i >> 2) & 1) ^ ((i >> 3) & 1), cc = ((i >> 1) & 1) ^ ((i >> 2) & 1),
                  d = (i & 1) ^ ((i >> 1) & 1);
        const int idx = ((d << 3) | (cc << 2) | (b << 1) | a) >> (4 - n);
        T[i] = (i & 1) ? -c[idx] : c[idx];
    }
}
// inverse ADST4 (spec 7.13.2.6)
inline void iadst4(int32_t *T, int r, uint32_t *err) {
    const int64_t S1 = 1321, S2 = 2482, S3 = 3344, S4 = 3803;
    int64_t s0 = S1 * T[0], s1 = S2 * T[0], s2 = S3 * T[1], s3 = S4 * T[2], s4 = S1 * T[2], s5 = S2 * T[3], s6 = S4 * T[3];
    const int64_t a7 = (int64_t)T[0] - T[2], b7 = a7 + T[3];
    chk(a7, r + 1, err, ERR_ADST4_RANGE);
    chk(b7, r, err, ERR_ADST4_RANGE);
    const int64_t sv[7] = {s0, s1, s2, s3, s4, s5, s6};
    for (int i = 0; i < 7; i++) chk(sv[i], r + 12, err, ERR_ADST4_RANGE);
    s0 = s0 + s3;
    s1 = s1 - s4;
    s3 = s2;
    s2 = S3 * b7;
    chk(s0, r + 12, err, ERR_ADST4_RANGE); chk(s1, r + 12, err, ERR_ADST4_RANGE); chk(s2, r + 12, err, ERR_ADST4_RANGE);
    s0 = s0 + s5;
    s1 = s1 - s6;
    chk(s0, r + 12, err, ERR_ADST4_RANGE); chk(s1, r + 12, err, ERR_ADST4_RANGE);
    int64_t x[4] = {s0 + s3, s1 + s3, s2, s0 + s1};
    chk(x[3], r + 12, err, ERR_ADST4_RANGE);
    x[3] = x[3] - s3;
    for (int i = 0; i < 4; i++) chk(x[i], r + 12, err, ERR_ADST4_RANGE);
    for (int i = 0; i < 4; i++) T[i] = (int32_t)round2_64(x[i], 12);
}
// inverse ADST8 / ADST16 (spec 7.13.2.7 / 7.13.2.8)
inline void iadst8(int32_t *T, int r, uint32_t *err) {
    adst_in_perm(T, 3);
    for (int i = 0; i < 4; i++) bfly(T, 2 * i, 2 * i + 1, 60 - 16 * i, 1, r, err);
    for (int i = 0; i < 4; i++) hbfly(T, i, 4 + i, 0, r);
    for (int i = 0; i < 2; i++) bfly(T, 4 + 3 * i, 5 + i, 48 - 32 * i, 1, r, err);
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) hbfly(T, 4 * j + i, 2 + 4 * j + i, 0, r);
    for (int i = 0; i < 2; i++) bfly(T, 2 + 4 * i, 3 + 4 * i, 32, 1, r, err);
    adst_out_perm(T, 3);
}
inline void iadst16(int32_t *T, int r, uint32_t *err) {
    adst_in_perm(T, 4);
    for (int i = 0; i < 8; i++) bfly(T, 2 * i, 2 * i + 1, 62 - 8 * i, 1, r, err);
    for (int i = 0; i < 8; i++) hbfly(T, i, 8 + i, 0, r);
    for (int i = 0; i < 2; i++) { bfly(T, 8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1, r, err); bfly(T, 13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1, r, err); }
    for (int i = 0; i < 4; i++) for (int j = 0; j < 2; j++) hbfly(T, 8 * j + i, 4 + 8 * j + i, 0, r);
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) bfly(T, 4 + 8 * j + 3 * i, 5 + 8 * j + i, 48 - 32 * i, 1, r, err);
    for (int i = 0; i < 2; i++) for (int j = 0; j < 4; j++) hbfly(T, 4 * j + i, 2 + 4 * j + i, 0, r);
    for (int i = 0; i < 4; i++) bfly(T, 2 + 4 * i, 3 + 4 * i, 32, 1, r, err);
    adst_out_perm(T, 4);
}
// inverse identity (spec 7.13.2.15)
inline void iidentity(int32_t *T, int n) {
    for (int i = 0; i < (1 << n); i++) {
        if (n == 2) T[i] = (int32_t)round2_64((int64_t)T[i] * 5793, 12);
        else if (n == 3) T[i] = T[i] * 2;
        else if (n == 4) T[i] = (int32_t)round2_64((int64_t)T[i] * 11586, 12);
        else T[i] = T[i] * 4;
    }
}
inline void inv_1d(int32_t *T, int kind, int n, int r, uint32_t *err) {
    if (kind == K_DCT) idct(T, n, r, err);
    else if (kind == K_ADST) { if (n == 2) iadst4(T, r, err); else if (n == 3) iadst8(T, r, err); else iadst16(T, r, err); }
    else iidentity(T, n);
}
// 2-D inverse transform process (spec 7.13.3, Lossless = 0; dsp.h:419-449).
// dq: th x tw row-major dequantized coefficients; res: h x w residual.
inline void inv_txfm2d(const int32_t *dq, int txSz, int txType, int32_t *res, uint32_t *err, int bd) {
    const int lw = av1t::Tx_Width_Log2[txSz], lh = av1t::Tx_Height_Log2[txSz], w = 1 << lw, h = 1 << lh;
    const int tw = w > 32 ? 32 : w, th = h > 32 ? 32 : h;
    const int rowShift = av1t::Transform_Row_Shift[txSz];
    const int rowR = bd + 8, colR = bd + 6 > 16 ? bd + 6 : 16;
    const int rk = row_kind(txType), ck = col_kind(txType);
    const bool rect = lw - lh == 1 || lh - lw == 1;
    int32_t T[64];
    for (int i = 0; i < h; i++) {
        bool nz = false;
        if (i < th) for (int j = 0; j < tw; j++) nz |= dq[i * tw + j] != 0;
        if (!nz) { for (int j = 0; j < w; j++) res[i * w + j] = 0; continue; }
        for (int j = 0; j < w; j++) T[j] = j < tw ? dq[i * tw + j] : 0;
        if (rect) for (int j = 0; j < w; j++) T[j] = (int32_t)round2_64((int64_t)T[j] * 2896, 12);
        inv_1d(T, rk, lw, rowR, err);
        for (int j = 0; j < w; j++)
            res[i * w + j] = clip3(-(1 << (colR - 1)), (1 << (colR - 1)) - 1, (int32_t)round2_64(T[j], rowShift));
    }
    for (int j = 0; j < w; j++) {
        for (int i = 0; i < h; i++) T[i] = res[i * w + j];
        inv_1d(T, ck, lh, colR, err);
        for (int i = 0; i < h; i++) res[i * w + j] = (int32_t)round2_64(T[i], 4);
    }
}

// ---- intra prediction (spec 7.11.2; dsp.h:528-829) ----
// AboveRow[-1 .. w+h-1] / LeftCol[-1 .. w+h-1] from plane (stride), tx block at (x, y) (dsp.h:528-562)
inline void build_edges(const uint16_t *plane, int stride, int x, int y, int lw, int lh, int haveL, int haveA, int haveAR,
                        int haveBL, int maxX, int maxY, uint16_t *above, uint16_t *left, int bd) {
    const int w = 1 << lw, h = 1 << lh, n = w + h;
    if (!haveA && haveL) { for (int i = 0; i < n; i++) above[i] = plane[y * stride + x - 1]; }
    else if (!haveA && !haveL) { for (int i = 0; i < n; i++) above[i] = (uint16_t)((1 << (bd - 1)) - 1); }
    else {
        int lim = x + (haveAR ? 2 * w : w) - 1;
        if (lim > maxX) lim = maxX;
        for (int i = 0; i < n; i++) above[i] = plane[(y - 1) * stride + (x + i < lim ? x + i : lim)];
    }
    if (!haveL && haveA) { for (int i = 0; i < n; i++) left[i] = plane[(y - 1) * stride + x]; }
    else if (!haveL && !haveA) { for (int i = 0; i < n; i++) left[i] = (uint16_t)((1 << (bd - 1)) + 1); }
    else {
        int lim = y + (haveBL ? 2 * h : h) - 1;
        if (lim > maxY) lim = maxY;
        for (int i = 0; i < n; i++) left[i] = plane[(y + i < lim ? y + i : lim) * stride + x - 1];
    }
    uint16_t c;
    if (haveA && haveL) c = plane[(y - 1) * stride + x - 1];
    else if (haveA) c = plane[(y - 1) * stride + x];
    else if (haveL) c = plane[y * stride + x - 1];
    else c = (uint16_t)(1 << (bd - 1));
    above[-1] = c;
    left[-1] = c;
}
inline bool is_directional(int m) { return m >= V_PRED && m <= D67_PRED; }
inline const uint8_t *sm_weights(int l) {
    switch (l) {
    case 2: return av1t::Sm_Weights_Tx_4x4; case 3: return av1t::Sm_Weights_Tx_8x8; case 4: return av1t::Sm_Weights_Tx_16x16;
    case 5: return av1t::Sm_Weights_Tx_32x32; default: return av1t::Sm_Weights_Tx_64x64;
    }
}
// DC value (spec 7.11.2.5)
inline int dc_value(const uint16_t *above, const uint16_t *left, int haveL, int haveA, int lw, int lh, int bd) {
    const int w = 1 << lw, h = 1 << lh;
    int sum = 0;
    if (haveL && haveA) {
        for (int k = 0; k < h; k++) sum += left[k];
        for (int k = 0; k < w; k++) sum += above[k];
        return (sum + ((w + h) >> 1)) / (w + h);
    }
    if (haveL) { for (int k = 0; k < h; k++) sum += left[k]; return clip1((sum + (h >> 1)) >> lh, bd); }
    if (haveA) { for (int k = 0; k < w; k++) sum += above[k]; return clip1((sum + (w >> 1)) >> lw, bd); }
    return 1 << (bd - 1);
}
inline int pangle_of(int mode, int delta) { return is_directional(mode) ? av1t::Mode_To_Angle[mode] + 3 * delta : 0; }
// one predicted sample of a non-DC mode at row i, column j (spec 7.11.2.2/4/6; dsp.h:602-663)
inline int pred_pixel(int mode, int pAngle, int i, int j, int lw, int lh, const uint16_t *above, const uint16_t *left,
                      int upA, int upL) {
    const int w = 1 << lw, h = 1 << lh;
    if (is_directional(mode)) {
        if (pAngle == 90) return above[j];
        if (pAngle == 180) return left[i];
        if (pAngle < 90) {
            const int dx = av1t::Dr_Intra_Derivative[pAngle], idx = (i + 1) * dx;
            const int base = (idx >> (6 - upA)) + (j << upA), shift = ((idx << upA) >> 1) & 0x1F;
            const int maxBaseX = (w + h - 1) << upA;
            if (base < maxBaseX) return round2(above[base] * (32 - shift) + above[base + 1] * shift, 5);
            return above[maxBaseX];
        }
        if (pAngle < 180) {
            const int dx = av1t::Dr_Intra_Derivative[180 - pAngle], dy = av1t::Dr_Intra_Derivative[pAngle - 90];
            int idx = (j << 6) - (i + 1) * dx, base = idx >> (6 - upA);
            if (base >= -(1 << upA)) {
                const int shift = ((idx * (1 << upA)) >> 1) & 0x1F;
                return round2(above[base] * (32 - shift) + above[base + 1] * shift, 5);
            }
            idx = (i << 6) - (j + 1) * dy;
            base = idx >> (6 - upL);
            const int shift = ((idx * (1 << upL)) >> 1) & 0x1F;
            return round2(left[base] * (32 - shift) + left[base + 1] * shift, 5);
        }
        const int dy = av1t::Dr_Intra_Derivative[270 - pAngle], idx = (j + 1) * dy;
        const int base = (idx >> (6 - upL)) + (i << upL), shift = ((idx << upL) >> 1) & 0x1F;
        return round2(left[base] * (32 - shift) + left[base + 1] * shift, 5);
    }
    if (mode == SMOOTH_PRED) {
        const uint8_t *wx = sm_weights(lw), *wy = sm_weights(lh);
        return round2(wy[i] * above[j] + (256 - wy[i]) * left[h - 1] + wx[j] * left[i] + (256 - wx[j]) * above[w - 1], 9);
    }
    if (mode == SMOOTH_V_PRED) {
        const uint8_t *wy = sm_weights(lh);
        return round2(wy[i] * above[j] + (256 - wy[i]) * left[h - 1], 8);
    }
    if (mode == SMOOTH_H_PRED) {
        const uint8_t *wx = sm_weights(lw);
        return round2(wx[j] * left[i] + (256 - wx[j]) * above[w - 1], 8);
    }
    const int base = above[j] + left[i] - above[-1];   // PAETH
    const int pL = base - left[i] < 0 ? left[i] - base : base - left[i];
    const int pT = base - above[j] < 0 ? above[j] - base : base - above[j];
    const int pTL = base - above[-1] < 0 ? above[-1] - base : base - above[-1];
    if (pL <= pT && pL <= pTL) return left[i];
    if (pT <= pTL) return above[j];
    return above[-1];
}
// intra edge filter strength / upsampling decisions (spec 7.11.2.9 / 7.11.2.10; dsp.h:666-691)
inline int edge_strength(int w, int h, int filterType, int delta) {
    const int d = delta < 0 ? -delta : delta, s = w + h;
    int st = 0;
    if (filterType == 0) {
        if (s <= 8) { if (d >= 56) st = 1; }
        else if (s <= 12) { if (d >= 40) st = 1; }
        else if (s <= 16) { if (d >= 40) st = 1; }
        else if (s <= 24) { if (d >= 8) st = 1; if (d >= 16) st = 2; if (d >= 32) st = 3; }
        else if (s <= 32) { st = 1; if (d >= 4) st = 2; if (d >= 32) st = 3; }
        else st = 3;
    } else {
        if (s <= 8) { if (d >= 40) st = 1; if (d >= 64) st = 2; }
        else if (s <= 16) { if (d >= 20) st = 1; if (d >= 48) st = 2; }
        else if (s <= 24) { if (d >= 4) st = 3; }
        else st = 3;
    }
    return st;
}
inline int edge_upsample(int w, int h, int filterType, int delta) {
    const int d = delta < 0 ? -delta : delta;
    if (d <= 0 || d >= 40) return 0;
    return filterType == 0 ? (w + h <= 16) : (w + h <= 8);
}
// intra edge filter of buf[0 .. sz-2] (spec 7.11.2.12)
inline void edge_filter(uint16_t *buf, int sz, int strength) {
    if (!strength) return;
    uint16_t e[160];
    for (int i = 0; i < sz; i++) e[i] = buf[i - 1];
    for (int i = 1; i < sz; i++) {
        int s = 0;
        for (int j = 0; j < 5; j++) s += av1t::Intra_Edge_Kernel[strength - 1][j] * e[clip3(0, sz - 1, i - 2 + j)];
        buf[i - 1] = (uint16_t)((s + 8) >> 4);
    }
}
// intra edge upsampling of buf[-1 .. numPx-1] into buf[-2 .. 2 numPx - 2] (spec 7.11.2.11)
inline void edge_upsample_px(uint16_t *buf, int numPx, int bd) {
    uint16_t dup[40];
    dup[0] = buf[-1];
    for (int i = -1; i < numPx; i++) dup[i + 2] = buf[i];
    dup[numPx + 2] = buf[numPx - 1];
    buf[-2] = dup[0];
    for (int i = 0; i < numPx; i++) {
        const int s = -dup[i] + 9 * dup[i + 1] + 9 * dup[i + 2] - dup[i + 3];
        buf[2 * i - 1] = (uint16_t)clip1(round2(s, 4), bd);
        buf[2 * i] = dup[i + 2];
    }
}
// edge preparation of a directional prediction with enable_intra_edge_filter (spec 7.11.2.4 step 4; dsp.h:723-751)
inline void edge_prepare(int pAngle, int lw, int lh, int filterType, int haveL, int haveA, int x, int y, int maxX, int maxY,
                         uint16_t *above, uint16_t *left, int *upA, int *upL, int bd) {
    const int w = 1 << lw, h = 1 << lh;
    *upA = *upL = 0;
    if (pAngle != 90 && pAngle != 180) {
        if (pAngle > 90 && pAngle < 180 && w + h >= 24) {
            const uint16_t c = (uint16_t)round2(left[0] * 5 + above[-1] * 6 + above[0] * 5, 4);
            left[-1] = c;
            above[-1] = c;
        }
        if (haveA)
            edge_filter(above, (w < maxX - x + 1 ? w : maxX - x + 1) + (pAngle < 90 ? h : 0) + 1, edge_strength(w, h, filterType, pAngle - 90));
        if (haveL)
            edge_filter(left, (h < maxY - y + 1 ? h : maxY - y + 1) + (pAngle > 180 ? w : 0) + 1, edge_strength(w, h, filterType, pAngle - 180));
    }
    *upA = edge_upsample(w, h, filterType, pAngle - 90);
    if (*upA) edge_upsample_px(above, w + (pAngle < 90 ? h : 0), bd);
    *upL = edge_upsample(w, h, filterType, pAngle - 180);
    if (*upL) edge_upsample_px(left, h + (pAngle > 180 ? w : 0), bd);
}
// Whole-tx-block prediction into pred[h][w] (row-major) of mode (not CfL, not filter intra), with the intra edge
// filter when edgeFilter (spec 7.11.2; dsp.h:786-804).  above / left are copied, the caller's stay unfiltered.
inline void predict_block(int mode, int delta, int edgeFilter, int filterType, int haveL, int haveA, int lw, int lh, int x,
                          int y, int maxX, int maxY, const uint16_t *above, const uint16_t *left, int dcv, int *pred, int bd) {
    const int w = 1 << lw, h = 1 << lh;
    if (mode == DC_PRED) { for (int k = 0; k < w * h; k++) pred[k] = dcv; return; }
    uint16_t ea[EDGE_LEN], el[EDGE_LEN];
    memcpy(ea, above - EDGE_OFF, sizeof ea);
    memcpy(el, left - EDGE_OFF, sizeof el);
    uint16_t *A = ea + EDGE_OFF, *L = el + EDGE_OFF;
    const int pa = pangle_of(mode, delta);
    int upA = 0, upL = 0;
    if (edgeFilter && is_directional(mode)) edge_prepare(pa, lw, lh, filterType, haveL, haveA, x, y, maxX, maxY, A, L, &upA, &upL, bd);
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) pred[i * w + j] = pred_pixel(mode, pa, i, j, lw, lh, A, L, upA, upL);
}

// ---- block geometry (blockinfo.h) ----
inline int has_chroma(int bs, int mi_row, int mi_col, int mono, int ssx, int ssy) {   // blockinfo.h:34-39
    if (mono) return 0;
    if (ssy && bh4(bs) == 1 && (mi_row & 1) == 0) return 0;
    if (ssx && bw4(bs) == 1 && (mi_col & 1) == 0) return 0;
    return 1;
}
inline int uv_bsize(int bs, int ssx, int ssy) { return av1t::Subsampled_Size[bs][ssx][ssy]; }
inline int uv_tx_size(int bs, int ssx, int ssy) {   // blockinfo.h:58-67
    const int t = av1t::Max_Tx_Size_Rect[uv_bsize(bs, ssx, ssy)];
    const int w = txw(t), h = txh(t);
    if (w == 64 || h == 64) { if (w == 16) return TX_16X32; if (h == 16) return TX_32X16; return TX_32X32; }
    return t;
}
inline int cfl_allowed(int bs) { return bw4(bs) <= 8 && bh4(bs) <= 8; }
// get_filter_type (spec 7.11.2.8; blockinfo.h:130-153)
inline int filter_type(int plane, int mi_row, int mi_col, int availU, int availL, const uint8_t *yModes, const uint8_t *uvModes,
                       int miCols, int ssx, int ssy) {
    auto smooth = [](int m) { return m == SMOOTH_PRED || m == SMOOTH_V_PRED || m == SMOOTH_H_PRED; };
    const uint8_t *m = plane ? uvModes : yModes;
    int a = 0, l = 0;
    if (availU) {
        int r = mi_row - 1, c = mi_col;
        if (plane > 0) { if (ssx && !(mi_col & 1)) c++; if (ssy && (mi_row & 1)) r--; }
        a = smooth(m[(size_t)r * miCols + c]);
    }
    if (availL) {
        int r = mi_row, c = mi_col - 1;
        if (plane > 0) { if (ssx && (mi_col & 1)) c--; if (ssy && !(mi_row & 1)) r++; }
        l = smooth(m[(size_t)r * miCols + c]);
    }
    return a || l;
}

}  // namespace orc
