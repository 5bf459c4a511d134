// Learned candidate pruning: a small BF16 CNN per 64x64 superblock, run on MFMA with FP32 accumulation, that predicts
// for each square partition node whether its children, its own shapes and its 4-way shapes are needed, and which luma
// modes to screen.  No counterpart in the reference (its pruning stages are hand-written heuristics:
// encoder/src/k_leafprep.cu:1-14, an_stage.h:1-15, the screener hook gp.h:136-143).  The network only removes
// candidates from the exhaustive search, so every decision it leads to is still an ordinary, decodable AV1 decision.
//
// Network (per superblock):  x0 = [64 x 64 luma / max - 0.5, qindex / 255]  (NHWC, 2 channels, BF16)
//   L1..L4: 3x3 convolution, stride 2, zero padding 1, 32 output channels, bias, ReLU   (64 -> 32 -> 16 -> 8 -> 4)
//   L5: fully connected 512 -> 268 logits: 85 "expand" logits (node 0: the 64x64 node, 1-4: 32x32 nodes, 5-20: 16x16,
//       21-84: 8x8, raster order per level: the node needs its children), 85 "own" logits (same nodes: it needs its
//       N x N, N x N/2 or N/2 x N shapes), 85 "four" logits (it needs its N x N/4 or N/4 x N shapes: HORZ_4 / VERT_4)
//       and 13 luma-mode logits (the mode is used by some block of the superblock).
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16_bf16 (positions x input taps times taps x channels), weights
// and activations BF16, accumulation and biases FP32.  Training runs the same forward kernels, then FP32 backward
// kernels (convolution weight gradients accumulated with float atomics) and Adam on FP32 master weights.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include "dev_eval.h"

namespace av1m {

typedef uint16_t bf16_t;
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int LN_NODES = 85, LN_OWN0 = 85, LN_FOUR0 = 170, LN_MODE0 = 255, LN_OUT = 268, LN_FEAT = 512, LN_C = 32;
constexpr int LN_MASKW = 12;   // mask words per superblock

__device__ __forceinline__ bf16_t to_bf16(float f) {   // round to nearest even
    const uint32_t u = __float_as_uint(f);
    return (bf16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---- input tiles: 64 x 64 luma of each superblock (padded planes) and the qindex plane ----
// superblocks s0 .. s0 + n - 1 of the batch go to slots 0 .. n - 1
__global__ void __launch_bounds__(256) k_ln_gather(const Img *imgs, const uint32_t *sbImg, uint32_t s0, uint32_t n, bf16_t *x0) {
    const uint32_t slot = blockIdx.x, sb = s0 + slot;
    if (slot >= n) return;
    const Img &I = imgs[sbImg[sb]];
    const uint32_t local = sb - I.sbBase;
    const int sx = (int)(local % I.sbCols) * 64, sy = (int)(local / I.sbCols) * 64;
    const float inv = 1.0f / (float)((1 << I.bd) - 1), q = (float)I.qindex / 255.0f;
    bf16_t *o = x0 + (size_t)slot * 64 * 64 * 2;
    for (int p = threadIdx.x; p < 64 * 64; p += 256) {
        const int y = p >> 6, x = p & 63;
        const size_t off = (size_t)(sy + y) * I.spitch[0] + sx + x;
        const int v = I.bd == 8 ? ((const uint8_t *)I.src[0])[off] : ((const uint16_t *)I.src[0])[off];
        o[2 * p] = to_bf16((float)v * inv - 0.5f);
        o[2 * p + 1] = to_bf16(q);
    }
}

// ---- input tiles from 64 x 64 8-bit luma tiles (training data) and their qindex ----
__global__ void __launch_bounds__(256) k_ln_tiles(const uint8_t *tiles, const uint8_t *q, uint32_t n, bf16_t *x0) {
    const uint32_t it = blockIdx.x;
    if (it >= n) return;
    const float qq = (float)q[it] / 255.0f;
    for (int p = threadIdx.x; p < 64 * 64; p += 256) {
        x0[((size_t)it * 4096 + p) * 2] = to_bf16((float)tiles[(size_t)it * 4096 + p] * (1.0f / 255.0f) - 0.5f);
        x0[((size_t)it * 4096 + p) * 2 + 1] = to_bf16(qq);
    }
}

// ---- convolution forward on MFMA: one wave computes 16 output positions x 32 channels of one superblock ----
// A (16 x 16): lane l holds positions l % 16, taps kb + 4 (l / 16) + 0..3; B (16 x 16): taps likewise, channel l % 16;
// D: lane l holds channel l % 16 of positions 4 (l / 16) + 0..3.
// Geometry of a 3x3 stride-2 convolution layer (zero padding) over HIN x HIN x CIN NHWC activations: 16-position
// tiles of the HOUT x HOUT output, K = 9 CIN taps in KB blocks of 16 (the weight rows are zero-padded to KP).
template <int CIN, int HIN>
struct LnConv {
    static constexpr int HOUT = HIN / 2, NPOS = HOUT * HOUT, K = 9 * CIN, KP = (K + 15) / 16 * 16, TILES = NPOS / 16, KB = KP / 16;
};

// the weights of every 16-tap block for lane l (4 consecutive taps of output channels l & 15 and 16 + (l & 15))
template <int CIN, int HIN>
__device__ __forceinline__ void conv_weights(const bf16_t *W, int l, short4_t (&b0)[LnConv<CIN, HIN>::KB], short4_t (&b1)[LnConv<CIN, HIN>::KB]) {
    using G = LnConv<CIN, HIN>;
    const int c = l & 15, kq = l >> 4;
#pragma unroll
    for (int kb = 0; kb < G::KB; kb++) {
        b0[kb] = *(const short4_t *)(W + (size_t)c * G::KP + 16 * kb + 4 * kq);
        b1[kb] = *(const short4_t *)(W + (size_t)(16 + c) * G::KP + 16 * kb + 4 * kq);
    }
}

// One tile: 16 output positions x 32 channels as two 16x16x16 BF16 MFMA accumulations over the tap blocks in
// ascending order; lane l carries position tile * 16 + (l & 15) of the A operand and taps 4 (l >> 4) .. + 3 of each
// block.  x is the item's input, in global memory or LDS.
template <int CIN, int HIN>
__device__ __forceinline__ void conv_tile(const bf16_t *x, const short4_t (&b0)[LnConv<CIN, HIN>::KB],
                                          const short4_t (&b1)[LnConv<CIN, HIN>::KB], int tile, int l, float4_t &acc0, float4_t &acc1) {
    using G = LnConv<CIN, HIN>;
    constexpr int K = G::K;
    const int row = l & 15, kq = l >> 4, pos = tile * 16 + row, oy = pos / G::HOUT, ox = pos % G::HOUT;
    acc0 = float4_t{0.f, 0.f, 0.f, 0.f};
    acc1 = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kbi = 0; kbi < G::KB; kbi++) {
        const int kb = 16 * kbi;
        short4_t a = {0, 0, 0, 0};
        if constexpr (CIN % 4 == 0) {
            // the lane's 4 taps are 4 consecutive channels of one input sample: one 8-byte load (NHWC, channel
            // offset a multiple of 4, buffers 256-byte aligned)
            const int k = kb + 4 * kq, t = k / CIN, ci = k % CIN, iy = 2 * oy - 1 + t / 3, ix = 2 * ox - 1 + t % 3;
            if (k < K && iy >= 0 && iy < HIN && ix >= 0 && ix < HIN) a = *(const short4_t *)(x + ((size_t)iy * HIN + ix) * CIN + ci);
        } else if constexpr (CIN == 2) {
            // two taps of two channels each: one 4-byte load per tap
            typedef short short2_t __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int k = kb + 4 * kq + 2 * h, t = k / 2, iy = 2 * oy - 1 + t / 3, ix = 2 * ox - 1 + t % 3;
                short2_t v = {0, 0};
                if (k < K && iy >= 0 && iy < HIN && ix >= 0 && ix < HIN) v = *(const short2_t *)(x + ((size_t)iy * HIN + ix) * 2);
                a[2 * h] = v[0];
                a[2 * h + 1] = v[1];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = kb + 4 * kq + j;
                bf16_t v = 0;
                if (k < K) {
                    const int t = k / CIN, ci = k % CIN, iy = 2 * oy - 1 + t / 3, ix = 2 * ox - 1 + t % 3;
                    if (iy >= 0 && iy < HIN && ix >= 0 && ix < HIN) v = x[((size_t)iy * HIN + ix) * CIN + ci];
                }
                a[j] = (short)v;
            }
        }
        acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0[kbi], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1[kbi], acc1, 0, 0, 0);
    }
}

// a tile's outputs (bias, ReLU, BF16) at their NHWC places in dst, the 16 positions' 32 channels from tile * 512
__device__ __forceinline__ void conv_tile_out(bf16_t *dst, int tile, int l, const float4_t &acc0, const float4_t &acc1, float bias0,
                                              float bias1) {
    const int c = l & 15, kq = l >> 4;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int p = tile * 16 + 4 * kq + r;
        dst[(size_t)p * LN_C + c] = to_bf16(fmaxf(acc0[r] + bias0, 0.f));
        dst[(size_t)p * LN_C + 16 + c] = to_bf16(fmaxf(acc1[r] + bias1, 0.f));
    }
}

// One wave computes TPW consecutive tiles of one item, the weights in registers for all of them (loaded once per
// wave); each tile's outputs (1 KB of consecutive NHWC activations) are staged in LDS and written with one 16-byte
// store per lane.
template <int CIN, int HIN, int TPW>
__global__ void __launch_bounds__(64) k_ln_conv(const bf16_t *in, const bf16_t *W, const float *bias, bf16_t *out, uint32_t n) {
    using G = LnConv<CIN, HIN>;
    constexpr int GROUPS = G::TILES / TPW;
    static_assert(G::TILES % TPW == 0, "whole groups of tiles");
    const uint32_t item = blockIdx.x / GROUPS, grp = blockIdx.x % GROUPS;
    if (item >= n) return;
    const int l = threadIdx.x, c = l & 15;
    const bf16_t *x = in + (size_t)item * HIN * HIN * CIN;
    short4_t b0[G::KB], b1[G::KB];
    conv_weights<CIN, HIN>(W, l, b0, b1);
    const float bias0 = bias[c], bias1 = bias[16 + c];
    bf16_t *o = out + (size_t)item * G::NPOS * LN_C;
    __shared__ __attribute__((aligned(16))) bf16_t stage[16 * LN_C];
    for (int tt = 0; tt < TPW; tt++) {
        const int tile = (int)grp * TPW + tt;
        float4_t acc0, acc1;
        conv_tile<CIN, HIN>(x, b0, b1, tile, l, acc0, acc1);
        conv_tile_out(stage, 0, l, acc0, acc1, bias0, bias1);
        __syncthreads();
        *(uint4 *)(o + (size_t)tile * 16 * LN_C + 8 * l) = *(const uint4 *)(stage + 8 * l);
        __syncthreads();
    }
}

// The first two layers for one item per 256-thread workgroup, the first layer's 32 x 32 x 32 output (64 KB) kept in
// LDS: the four waves take 16 of its 64 tiles each, then 4 of the second layer's 16 tiles each, reading it from LDS.
// The tiles' arithmetic is conv_tile's, so the second layer's output equals k_ln_conv<2, 64> then k_ln_conv<32, 32>.
__global__ void __launch_bounds__(256) k_ln_conv12(const bf16_t *in, const bf16_t *W1, const float *bias1, const bf16_t *W2,
                                                   const float *bias2, bf16_t *out, uint32_t n) {
    using G1 = LnConv<2, 64>;
    using G2 = LnConv<32, 32>;
    const uint32_t item = blockIdx.x;
    if (item >= n) return;
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, c = l & 15;
    __shared__ __attribute__((aligned(16))) bf16_t a1[G1::NPOS * LN_C];
    __shared__ __attribute__((aligned(16))) bf16_t stage[4][16 * LN_C];
    {
        const bf16_t *x = in + (size_t)item * 64 * 64 * 2;
        short4_t b0[G1::KB], b1[G1::KB];
        conv_weights<2, 64>(W1, l, b0, b1);
        const float bb0 = bias1[c], bb1 = bias1[16 + c];
        for (int tt = 0; tt < G1::TILES / 4; tt++) {
            const int tile = w * (G1::TILES / 4) + tt;
            float4_t acc0, acc1;
            conv_tile<2, 64>(x, b0, b1, tile, l, acc0, acc1);
            conv_tile_out(a1, tile, l, acc0, acc1, bb0, bb1);
        }
    }
    __syncthreads();
    short4_t b0[G2::KB], b1[G2::KB];
    conv_weights<32, 32>(W2, l, b0, b1);
    const float bb0 = bias2[c], bb1 = bias2[16 + c];
    bf16_t *o = out + (size_t)item * G2::NPOS * LN_C;
    for (int tt = 0; tt < G2::TILES / 4; tt++) {
        const int tile = w * (G2::TILES / 4) + tt;
        float4_t acc0, acc1;
        conv_tile<32, 32>(a1, b0, b1, tile, l, acc0, acc1);
        conv_tile_out(stage[w], 0, l, acc0, acc1, bb0, bb1);
        __syncthreads();
        *(uint4 *)(o + (size_t)tile * 16 * LN_C + 8 * l) = *(const uint4 *)(stage[w] + 8 * l);
        __syncthreads();
    }
}

// ---- fully connected head: logits[item][o] = b[o] + sum_f W[o][f] a4[item][f] (FP32, f ascending) ----
// LN_FC_IT items per block share each weight read; their features sit in LDS (one fmaf chain per item and output)
constexpr int LN_FC_IT = 16;
__global__ void __launch_bounds__(128) k_ln_fc(const bf16_t *a4, const float *W, const float *b, float *logits, uint32_t n) {
    __shared__ float f[LN_FC_IT][LN_FEAT];
    const uint32_t i0 = blockIdx.x * LN_FC_IT;
    if (i0 >= n) return;
    const int ni = (int)min((uint32_t)LN_FC_IT, n - i0);
    for (int e = threadIdx.x; e < LN_FC_IT * LN_FEAT; e += 128) {
        const int i = e / LN_FEAT, k = e % LN_FEAT;
        f[i][k] = i < ni ? from_bf16(a4[(size_t)(i0 + i) * LN_FEAT + k]) : 0.f;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < LN_OUT; o += 128) {
        const float *w = W + (size_t)o * LN_FEAT;
        float s[LN_FC_IT];
#pragma unroll
        for (int i = 0; i < LN_FC_IT; i++) s[i] = b[o];
        for (int k = 0; k < LN_FEAT; k++) {
            const float wk = w[k];
#pragma unroll
            for (int i = 0; i < LN_FC_IT; i++) s[i] = fmaf(wk, f[i][k], s[i]);
        }
#pragma unroll
        for (int i = 0; i < LN_FC_IT; i++)
            if (i < ni) logits[(size_t)(i0 + i) * LN_OUT + o] = s[i];
    }
}

// ---- decisions: per superblock the nodes whose children / own / 4-way shapes are analysed, the luma modes screened ----
// A node is expanded (its children analysed) unless the network is confident (p <= 1 - conf) that the search does not
// need its children; likewise it keeps its own (N x N, N x N/2, N/2 x N) and its 4-way (N x N/4, N/4 x N) shapes.  A
// node where all three would be dropped keeps its own shapes (every node reached keeps a partition), and a node reaching
// past the frame keeps everything (the partition syntax may force a split there).  The modes are the 8 most likely
// when the network is confident that none of the other 5 is used (product of their 1 - p at least conf), else all 13.
// mask (12 words per superblock): 0-2 expand bits, 3 screened modes, 4-6 own bits, 8-10 four bits, 7 and 11 zero.
// logits of slots 0 .. n - 1 belong to superblocks s0 .. s0 + n - 1 of the batch; mask is indexed by superblock
__global__ void __launch_bounds__(64) k_ln_decide(const float *logits, const Img *imgs, const uint32_t *sbImg, uint32_t s0, uint32_t n,
                                                  float conf, uint32_t *mask) {
    const uint32_t slot = blockIdx.x * 64 + threadIdx.x, sb = s0 + slot;
    if (slot >= n) return;
    const Img &I = imgs[sbImg[sb]];
    const uint32_t local = sb - I.sbBase;
    const int sx = (int)(local % I.sbCols) * 64, sy = (int)(local / I.sbCols) * 64;
    const int FW = I.MiCols * 4, FH = I.MiRows * 4;
    const float *z = logits + (size_t)slot * LN_OUT;
    uint32_t m[LN_MASKW] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int node = 0;
    for (int lv = 0; lv < 4; lv++) {
        const int n = 64 >> lv, per = 1 << lv;
        for (int ny = 0; ny < per; ny++)
            for (int nx = 0; nx < per; nx++, node++) {
                const bool edge = sx + nx * n + n > FW || sy + ny * n + n > FH;
                const float pe = 1.f / (1.f + expf(-z[node])), po = 1.f / (1.f + expf(-z[LN_OWN0 + node]));
                const float pf = 1.f / (1.f + expf(-z[LN_FOUR0 + node]));
                const bool ex = edge || pe > 1.f - conf, fo = edge || pf > 1.f - conf;
                const bool ow = edge || po > 1.f - conf || (!ex && !fo);
                if (ex) m[node >> 5] |= 1u << (node & 31);
                if (ow) m[4 + (node >> 5)] |= 1u << (node & 31);
                if (fo) m[8 + (node >> 5)] |= 1u << (node & 31);
            }
    }
    float p[13];
    for (int k = 0; k < 13; k++) p[k] = 1.f / (1.f + expf(-z[LN_MODE0 + k]));
    uint32_t top = 0;
    for (int r = 0; r < 8; r++) {
        int b = -1;
        for (int k = 0; k < 13; k++) if (!((top >> k) & 1) && (b < 0 || p[k] > p[b])) b = k;
        top |= 1u << b;
    }
    float none = 1.f;   // probability that no mode outside the shortlist is used
    for (int k = 0; k < 13; k++) if (!((top >> k) & 1)) none *= 1.f - p[k];
    m[3] = none >= conf ? top : 0x1FFFu;
    for (int k = 0; k < LN_MASKW; k++) mask[(size_t)sb * LN_MASKW + k] = m[k];
}

// ================================ training ================================

// loss (binary cross-entropy with logits, masked) and its gradient / count
__global__ void __launch_bounds__(128) k_ln_loss(const float *logits, const float *label, const float *lmask, uint32_t n, float scale,
                                                 float *dlog, float *lossOut) {
    const uint32_t i = blockIdx.x * 128 + threadIdx.x;
    float l = 0.f;
    if (i < n * LN_OUT) {
        const float z = logits[i], y = label[i], w = lmask[i];
        const float s = 1.f / (1.f + expf(-z));
        dlog[i] = w * (s - y) * scale;
        l = w * (fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z))));
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(lossOut, l * scale);
}

// FC backward, inputs: da4[item][f] = sum_o dlog[item][o] W[o][f], masked by the ReLU of a4
__global__ void __launch_bounds__(256) k_ln_fc_bwd_in(const bf16_t *a4, const float *W, const float *dlog, uint32_t n, float *da4) {
    const uint32_t item = blockIdx.x;
    if (item >= n) return;
    const float *g = dlog + (size_t)item * LN_OUT;
    for (int k = threadIdx.x; k < LN_FEAT; k += 256) {
        float s = 0.f;
        for (int o = 0; o < LN_OUT; o++) s = fmaf(g[o], W[(size_t)o * LN_FEAT + k], s);
        da4[(size_t)item * LN_FEAT + k] = from_bf16(a4[(size_t)item * LN_FEAT + k]) > 0.f ? s : 0.f;
    }
}
// FC backward, weights: one thread per (o, f) over the items (no atomics); the bias row uses f = LN_FEAT
__global__ void __launch_bounds__(256) k_ln_fc_bwd_w(const bf16_t *a4, const float *dlog, uint32_t n, float *dW, float *db) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= (uint32_t)LN_OUT * (LN_FEAT + 1)) return;
    const int o = (int)(e / (LN_FEAT + 1)), k = (int)(e % (LN_FEAT + 1));
    float s = 0.f;
    for (uint32_t it = 0; it < n; it++) {
        const float g = dlog[(size_t)it * LN_OUT + o];
        s = fmaf(g, k < LN_FEAT ? from_bf16(a4[(size_t)it * LN_FEAT + k]) : 1.f, s);
    }
    if (k < LN_FEAT) dW[(size_t)o * LN_FEAT + k] += s;
    else db[o] += s;
}

// convolution backward.  dz: gradient at the convolution's output before the ReLU (already masked), FP32 NHWC.
// Weights: dW[co][k] += sum over items and output positions of dz[pos][co] * im2col(in)[pos][k]; db[co] += sum dz.
template <int CIN, int HIN>
__global__ void __launch_bounds__(256) k_ln_conv_bwd_w(const bf16_t *in, const float *dz, uint32_t n, float *dW, float *db) {
    constexpr int HOUT = HIN / 2, NPOS = HOUT * HOUT, K = 9 * CIN, KP = (K + 15) / 16 * 16;
    const uint32_t item = blockIdx.x;
    if (item >= n) return;
    const bf16_t *x = in + (size_t)item * HIN * HIN * CIN;
    const float *g = dz + (size_t)item * NPOS * LN_C;
    for (int e = threadIdx.x; e < LN_C * K; e += 256) {
        const int co = e / K, k = e % K, t = k / CIN, ci = k % CIN, ky = t / 3, kx = t % 3;
        float s = 0.f;
        for (int pos = 0; pos < NPOS; pos++) {
            const int oy = pos / HOUT, ox = pos % HOUT, iy = 2 * oy - 1 + ky, ix = 2 * ox - 1 + kx;
            if (iy < 0 || iy >= HIN || ix < 0 || ix >= HIN) continue;
            s = fmaf(g[(size_t)pos * LN_C + co], from_bf16(x[((size_t)iy * HIN + ix) * CIN + ci]), s);
        }
        if (s != 0.f) atomicAdd(&dW[(size_t)co * KP + k], s);
    }
    for (int co = threadIdx.x; co < LN_C; co += 256) {
        float s = 0.f;
        for (int pos = 0; pos < NPOS; pos++) s += g[(size_t)pos * LN_C + co];
        if (s != 0.f) atomicAdd(&db[co], s);
    }
}
// Input: dIn[iy][ix][ci] = sum over taps (ky, kx) and channels co of dz[oy][ox][co] W[co][tap][ci] with
// iy = 2 oy - 1 + ky, then masked by the ReLU of that input (the previous layer's output > 0): the previous dz.
template <int CIN, int HIN>
__global__ void __launch_bounds__(256) k_ln_conv_bwd_in(const bf16_t *in, const float *W, const float *dz, uint32_t n, float *dzPrev) {
    constexpr int HOUT = HIN / 2, K = 9 * CIN, KP = (K + 15) / 16 * 16;
    const uint32_t item = blockIdx.x;
    if (item >= n) return;
    const bf16_t *x = in + (size_t)item * HIN * HIN * CIN;
    const float *g = dz + (size_t)item * HOUT * HOUT * LN_C;
    float *d = dzPrev + (size_t)item * HIN * HIN * CIN;
    for (int e = threadIdx.x; e < HIN * HIN * CIN; e += 256) {
        const int ci = e % CIN, p = e / CIN, iy = p / HIN, ix = p % HIN;
        float s = 0.f;
        if (from_bf16(x[e]) > 0.f) {
            for (int ky = 0; ky < 3; ky++) {
                const int ty = iy + 1 - ky;
                if (ty < 0 || (ty & 1) || (ty >> 1) >= HOUT) continue;
                for (int kx = 0; kx < 3; kx++) {
                    const int tx = ix + 1 - kx;
                    if (tx < 0 || (tx & 1) || (tx >> 1) >= HOUT) continue;
                    const float *gz = g + ((size_t)(ty >> 1) * HOUT + (tx >> 1)) * LN_C;
                    const int k = (ky * 3 + kx) * CIN + ci;
                    for (int co = 0; co < LN_C; co++) s = fmaf(gz[co], W[(size_t)co * KP + k], s);
                }
            }
        }
        d[e] = s;
    }
}

// Adam on the FP32 master weights; refreshes the BF16 copy (bf may be null)
__global__ void __launch_bounds__(256) k_ln_adam(float *w, const float *g, float *m, float *v, bf16_t *bf, uint32_t n, float lr, float b1,
                                                 float b2, float eps, float c1, float c2, float wd) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    m[i] = b1 * m[i] + (1.f - b1) * gi;
    v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
    w[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
    if (bf) bf[i] = to_bf16(w[i]);
}

__global__ void __launch_bounds__(256) k_ln_to_bf16(const float *w, bf16_t *bf, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) bf[i] = to_bf16(w[i]);
}

// ================================ host side ================================

// parameter layout (FP32 masters): per convolution l = 1..4 the weights [32][KP_l] (taps past 9 Cin are zero) and the
// bias [32], then the FC weights [98][512] and bias [98]
constexpr int LN_KP1 = 32, LN_KP = 288;
constexpr size_t LN_W_OFF[4] = {0, 32 * LN_KP1 + 32, 32 * LN_KP1 + 32 + 32 * LN_KP + 32, 32 * LN_KP1 + 32 + 2 * (32 * LN_KP + 32)};
constexpr size_t LN_FC_OFF = 32 * LN_KP1 + 32 + 3 * (32 * LN_KP + 32);
constexpr size_t LN_NPARAM = LN_FC_OFF + (size_t)LN_OUT * LN_FEAT + LN_OUT;
constexpr size_t LN_NBF = 32 * LN_KP1 + 3 * 32 * LN_KP;   // the BF16 copies of the convolution weights
inline size_t ln_kp(int l) { return l == 0 ? LN_KP1 : LN_KP; }
inline size_t ln_bf_off(int l) { return l == 0 ? 0 : 32 * LN_KP1 + (size_t)(l - 1) * 32 * LN_KP; }

// activations of n items: the input and the four convolution outputs (BF16, NHWC) and the logits
struct LnActs {
    bf16_t *x0 = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *a4 = nullptr;
    float *logits = nullptr;
};
constexpr size_t LN_ACT_SZ[6] = {64 * 64 * 2 * 2, 32 * 32 * 32 * 2, 16 * 16 * 32 * 2, 8 * 8 * 32 * 2, LN_FEAT * 2, LN_OUT * 4};
inline size_t ln_al(size_t b) { return (b + 255) & ~(size_t)255; }
// device bytes of the activations of n items (one buffer, carved by ln_carve)
size_t ln_scratch_bytes(uint32_t n) {
    size_t t = 0;
    for (size_t b : LN_ACT_SZ) t += ln_al(b * n);
    return t;
}
static LnActs ln_carve(void *base, uint32_t n) {
    char *p = (char *)base;
    LnActs a;
    void **dst[6] = {(void **)&a.x0, (void **)&a.a1, (void **)&a.a2, (void **)&a.a3, (void **)&a.a4, (void **)&a.logits};
    for (int k = 0; k < 6; k++) { *dst[k] = p; p += ln_al(LN_ACT_SZ[k] * n); }
    return a;
}

struct LnNet {
    float *w = nullptr;            // FP32 parameters (LN_NPARAM)
    bf16_t *wb = nullptr;          // BF16 convolution weights (LN_NBF)
    // activations of the training and check entry points, for up to cap items (one allocation)
    uint32_t cap = 0;
    void *actMem = nullptr;
    LnActs act;
    // training state (allocated on first use)
    float *g = nullptr, *m = nullptr, *v = nullptr, *dlog = nullptr, *d4 = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr, *loss = nullptr;
    float *label = nullptr, *lmask = nullptr;
    uint8_t *tiles = nullptr, *q = nullptr;
    uint32_t dcap = 0;             // items the gradient buffers hold
    uint32_t ndata = 0;            // items of training data on the device
    int step = 0;
};

static hipError_t ln_alloc_acts(LnNet &N, uint32_t n) {
    if (n <= N.cap) return hipSuccess;
    (void)hipFree(N.actMem);
    N.actMem = nullptr; N.act = LnActs(); N.cap = 0;
    const hipError_t e = hipMalloc(&N.actMem, ln_scratch_bytes(n));
    if (e != hipSuccess) return e;
    N.act = ln_carve(N.actMem, n);
    N.cap = n;
    return hipSuccess;
}

// forward of n items whose inputs are in a.x0
// fused: the first two layers in one launch (k_ln_conv12) without writing the first layer's activations, for
// inference; training and the activation check read them
static void ln_forward(const LnNet &N, const LnActs &a, uint32_t n, hipStream_t st, bool fused = false) {
    const float *w = N.w;
    // grids: items x (tiles / tiles per wave); 64, 16, 4 and 1 tiles per item
    if (fused) {
        hipLaunchKernelGGL(k_ln_conv12, dim3(n), dim3(256), 0, st, a.x0, N.wb + ln_bf_off(0), w + LN_W_OFF[0] + 32 * LN_KP1, N.wb + ln_bf_off(1),
                           w + LN_W_OFF[1] + 32 * LN_KP, a.a2, n);
    } else {
        hipLaunchKernelGGL((k_ln_conv<2, 64, 8>), dim3(n * 8), dim3(64), 0, st, a.x0, N.wb + ln_bf_off(0), w + LN_W_OFF[0] + 32 * LN_KP1, a.a1, n);
        hipLaunchKernelGGL((k_ln_conv<32, 32, 4>), dim3(n * 4), dim3(64), 0, st, a.a1, N.wb + ln_bf_off(1), w + LN_W_OFF[1] + 32 * LN_KP, a.a2, n);
    }
    hipLaunchKernelGGL((k_ln_conv<32, 16, 4>), dim3(n), dim3(64), 0, st, a.a2, N.wb + ln_bf_off(2), w + LN_W_OFF[2] + 32 * LN_KP, a.a3, n);
    hipLaunchKernelGGL((k_ln_conv<32, 8, 1>), dim3(n), dim3(64), 0, st, a.a3, N.wb + ln_bf_off(3), w + LN_W_OFF[3] + 32 * LN_KP, a.a4, n);
    hipLaunchKernelGGL(k_ln_fc, dim3((n + LN_FC_IT - 1) / LN_FC_IT), dim3(128), 0, st, a.a4, w + LN_FC_OFF, w + LN_FC_OFF + (size_t)LN_OUT * LN_FEAT,
                       a.logits, n);
}

static void ln_refresh_bf16(LnNet &N, hipStream_t st) {
    for (int l = 0; l < 4; l++)
        hipLaunchKernelGGL(k_ln_to_bf16, dim3((unsigned)((32 * ln_kp(l) + 255) / 256)), dim3(256), 0, st, N.w + LN_W_OFF[l],
                           N.wb + ln_bf_off(l), (uint32_t)(32 * ln_kp(l)));
}

static hipError_t ln_upload(LnNet &N, const float *params) {
    hipError_t e;
    if (!N.w && (e = hipMalloc(&N.w, LN_NPARAM * 4)) != hipSuccess) return e;
    if (!N.wb && (e = hipMalloc(&N.wb, LN_NBF * 2)) != hipSuccess) return e;
    if ((e = hipMemcpy(N.w, params, LN_NPARAM * 4, hipMemcpyHostToDevice)) != hipSuccess) return e;
    ln_refresh_bf16(N, 0);
    return hipDeviceSynchronize();
}

static void ln_free(LnNet &N) {
    void *p[] = {N.w, N.wb, N.actMem, N.g, N.m, N.v, N.dlog, N.d4, N.d3, N.d2, N.d1, N.loss, N.label, N.lmask,
                 N.tiles, N.q};
    for (void *x : p) if (x) (void)hipFree(x);
    N = LnNet();
}

LnNet *ln_new(const float *params) {
    LnNet *N = new LnNet();
    if (ln_upload(*N, params) != hipSuccess) { ln_free(*N); delete N; return nullptr; }
    return N;
}
void ln_delete(LnNet *N) {
    if (!N) return;
    ln_free(*N);
    delete N;
}

// learned masks of a batch (encoder integration): mask[nSb][12].  scratch: ln_scratch_bytes(chunk) device bytes
// (stream-ordered memory of the encode call, so the network allocates and frees nothing itself).
int ln_masks(const LnNet &N, void *scratch, const Img *dImg, const uint32_t *dSbImg, uint32_t nSb, float conf, uint32_t chunk,
             uint32_t *dMask, hipStream_t st) {
    const LnActs a = ln_carve(scratch, chunk);
    for (uint32_t s0 = 0; s0 < nSb; s0 += chunk) {
        const uint32_t n = nSb - s0 < chunk ? nSb - s0 : chunk;
        hipLaunchKernelGGL(k_ln_gather, dim3(n), dim3(256), 0, st, dImg, dSbImg, s0, n, a.x0);
        ln_forward(N, a, n, st, true);
        hipLaunchKernelGGL(k_ln_decide, dim3((n + 63) / 64), dim3(64), 0, st, a.logits, dImg, dSbImg, s0, n, conf, dMask);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace av1m

// ================================ C ABI ================================
using namespace av1m;

struct av1mi_learner { LnNet net; };

extern "C" {

size_t av1mi_ln_param_count(void) { return LN_NPARAM; }

// He-normal initialisation from a seed (Box-Muller over splitmix64), zero biases and zero padding taps
void av1mi_ln_init_params(float *p, uint64_t seed) {
    auto next = [&]() {
        seed += 0x9E3779B97F4A7C15ull;
        uint64_t z = seed;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() { return sqrt(-2.0 * log(next() + 1e-300)) * cos(6.283185307179586 * next()); };
    memset(p, 0, LN_NPARAM * 4);
    for (int l = 0; l < 4; l++) {
        const int cin = l == 0 ? 2 : 32, K = 9 * cin;
        const double sd = sqrt(2.0 / K);
        for (int co = 0; co < 32; co++)
            for (int k = 0; k < K; k++) p[LN_W_OFF[l] + (size_t)co * ln_kp(l) + k] = (float)(sd * gauss());
    }
    const double sd = sqrt(1.0 / LN_FEAT);
    for (size_t i = 0; i < (size_t)LN_OUT * LN_FEAT; i++) p[LN_FC_OFF + i] = (float)(sd * gauss());
}

int av1mi_ln_create(const float *params, av1mi_learner **out) {
    av1mi_learner *L = new av1mi_learner();
    if (ln_upload(L->net, params) != hipSuccess) { ln_free(L->net); delete L; return -1; }
    *out = L;
    return 0;
}
void av1mi_ln_destroy(av1mi_learner *L) {
    if (!L) return;
    ln_free(L->net);
    delete L;
}
int av1mi_ln_get_params(av1mi_learner *L, float *params) {
    return hipMemcpy(params, L->net.w, LN_NPARAM * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

// forward of n 64x64 8-bit luma tiles (host arrays: tiles n x 4096, q n) -> logits (n x 98)
int av1mi_ln_forward(av1mi_learner *L, const uint8_t *tiles, const uint8_t *q, uint32_t n, float *logits) {
    LnNet &N = L->net;
    if (n == 0) return 0;
    if (ln_alloc_acts(N, n) != hipSuccess) return -1;
    uint8_t *dt = nullptr, *dq = nullptr;
    if (hipMalloc(&dt, (size_t)n * 4096) != hipSuccess || hipMalloc(&dq, n) != hipSuccess) { (void)hipFree(dt); return -1; }
    (void)hipMemcpy(dt, tiles, (size_t)n * 4096, hipMemcpyHostToDevice);
    (void)hipMemcpy(dq, q, n, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_ln_tiles, dim3(n), dim3(256), 0, 0, dt, dq, n, N.act.x0);
    ln_forward(N, N.act, n, 0);
    const hipError_t e = hipMemcpy(logits, N.act.logits, (size_t)n * LN_OUT * 4, hipMemcpyDeviceToHost);
    (void)hipFree(dt); (void)hipFree(dq);
    return e == hipSuccess ? 0 : -1;
}

// the activations of the last forward (for the checks against the CPU): layer 0 = x0 .. 4 = a4, as BF16 bits
int av1mi_ln_activations(av1mi_learner *L, int layer, uint32_t n, uint16_t *out) {
    LnNet &N = L->net;
    const size_t per[5] = {64 * 64 * 2, 32 * 32 * 32, 16 * 16 * 32, 8 * 8 * 32, LN_FEAT};
    const bf16_t *src[5] = {N.act.x0, N.act.a1, N.act.a2, N.act.a3, N.act.a4};
    if (layer < 0 || layer > 4 || n > N.cap) return -1;
    return hipMemcpy(out, src[layer], n * per[layer] * 2, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

// Training data on the device: n tiles (n x 4096 8-bit luma), their qindex, labels and label masks (n x 98 each).
int av1mi_ln_set_data(av1mi_learner *L, const uint8_t *tiles, const uint8_t *q, const float *label, const float *lmask, uint32_t n) {
    LnNet &N = L->net;
    (void)hipFree(N.tiles); (void)hipFree(N.q); (void)hipFree(N.label); (void)hipFree(N.lmask);
    N.tiles = N.q = nullptr; N.label = N.lmask = nullptr; N.ndata = 0;
    if (n == 0) return -1;
    if (hipMalloc(&N.tiles, (size_t)n * 4096) != hipSuccess || hipMalloc(&N.q, n) != hipSuccess ||
        hipMalloc(&N.label, (size_t)n * LN_OUT * 4) != hipSuccess || hipMalloc(&N.lmask, (size_t)n * LN_OUT * 4) != hipSuccess)
        return -1;
    (void)hipMemcpy(N.tiles, tiles, (size_t)n * 4096, hipMemcpyHostToDevice);
    (void)hipMemcpy(N.q, q, n, hipMemcpyHostToDevice);
    (void)hipMemcpy(N.label, label, (size_t)n * LN_OUT * 4, hipMemcpyHostToDevice);
    if (hipMemcpy(N.lmask, lmask, (size_t)n * LN_OUT * 4, hipMemcpyHostToDevice) != hipSuccess) return -1;
    N.ndata = n;
    return 0;
}

// Gradient of the mean masked BCE over items [first, first + n) of the data into N.g (and, with grads != null, to the
// host); returns the loss through *loss.
static int ln_grad(LnNet &N, uint32_t first, uint32_t n, float *loss, hipStream_t st) {
    if (n == 0 || first >= N.ndata || n > N.ndata - first) return -1;   // items past the data on the device
    if (ln_alloc_acts(N, n) != hipSuccess) return -1;
    if (!N.g) {
        if (hipMalloc(&N.g, LN_NPARAM * 4) != hipSuccess || hipMalloc(&N.m, LN_NPARAM * 4) != hipSuccess ||
            hipMalloc(&N.v, LN_NPARAM * 4) != hipSuccess || hipMalloc(&N.loss, 4) != hipSuccess)
            return -1;
        (void)hipMemset(N.m, 0, LN_NPARAM * 4);
        (void)hipMemset(N.v, 0, LN_NPARAM * 4);
    }
    if (n > N.dcap || !N.dlog) {
        (void)hipFree(N.dlog); (void)hipFree(N.d4); (void)hipFree(N.d3); (void)hipFree(N.d2); (void)hipFree(N.d1);
        if (hipMalloc(&N.dlog, (size_t)n * LN_OUT * 4) != hipSuccess || hipMalloc(&N.d4, (size_t)n * LN_FEAT * 4) != hipSuccess ||
            hipMalloc(&N.d3, (size_t)n * 8 * 8 * 32 * 4) != hipSuccess || hipMalloc(&N.d2, (size_t)n * 16 * 16 * 32 * 4) != hipSuccess ||
            hipMalloc(&N.d1, (size_t)n * 32 * 32 * 32 * 4) != hipSuccess)
            return -1;
        N.dcap = n;
    }
    (void)hipMemsetAsync(N.g, 0, LN_NPARAM * 4, st);
    (void)hipMemsetAsync(N.loss, 0, 4, st);
    hipLaunchKernelGGL(k_ln_tiles, dim3(n), dim3(256), 0, st, N.tiles + (size_t)first * 4096, N.q + first, n, N.act.x0);
    ln_forward(N, N.act, n, st);
    const float scale = 1.0f / (float)n;
    hipLaunchKernelGGL(k_ln_loss, dim3((n * LN_OUT + 127) / 128), dim3(128), 0, st, N.act.logits, N.label + (size_t)first * LN_OUT,
                       N.lmask + (size_t)first * LN_OUT, n, scale, N.dlog, N.loss);
    float *g = N.g;
    const float *w = N.w;
    hipLaunchKernelGGL(k_ln_fc_bwd_w, dim3((LN_OUT * (LN_FEAT + 1) + 255) / 256), dim3(256), 0, st, N.act.a4, N.dlog, n, g + LN_FC_OFF,
                       g + LN_FC_OFF + (size_t)LN_OUT * LN_FEAT);
    hipLaunchKernelGGL(k_ln_fc_bwd_in, dim3(n), dim3(256), 0, st, N.act.a4, w + LN_FC_OFF, N.dlog, n, N.d4);
    // layer 4 (input a3, 8x8): weights, then the gradient at a3 (masked by its ReLU)
    hipLaunchKernelGGL((k_ln_conv_bwd_w<32, 8>), dim3(n), dim3(256), 0, st, N.act.a3, N.d4, n, g + LN_W_OFF[3], g + LN_W_OFF[3] + 32 * LN_KP);
    hipLaunchKernelGGL((k_ln_conv_bwd_in<32, 8>), dim3(n), dim3(256), 0, st, N.act.a3, w + LN_W_OFF[3], N.d4, n, N.d3);
    hipLaunchKernelGGL((k_ln_conv_bwd_w<32, 16>), dim3(n), dim3(256), 0, st, N.act.a2, N.d3, n, g + LN_W_OFF[2], g + LN_W_OFF[2] + 32 * LN_KP);
    hipLaunchKernelGGL((k_ln_conv_bwd_in<32, 16>), dim3(n), dim3(256), 0, st, N.act.a2, w + LN_W_OFF[2], N.d3, n, N.d2);
    hipLaunchKernelGGL((k_ln_conv_bwd_w<32, 32>), dim3(n), dim3(256), 0, st, N.act.a1, N.d2, n, g + LN_W_OFF[1], g + LN_W_OFF[1] + 32 * LN_KP);
    hipLaunchKernelGGL((k_ln_conv_bwd_in<32, 32>), dim3(n), dim3(256), 0, st, N.act.a1, w + LN_W_OFF[1], N.d2, n, N.d1);
    hipLaunchKernelGGL((k_ln_conv_bwd_w<2, 64>), dim3(n), dim3(256), 0, st, N.act.x0, N.d1, n, g + LN_W_OFF[0], g + LN_W_OFF[0] + 32 * LN_KP1);
    if (hipGetLastError() != hipSuccess) return -1;
    if (loss && hipMemcpyAsync(loss, N.loss, 4, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

int av1mi_ln_gradient(av1mi_learner *L, uint32_t first, uint32_t n, float *loss, float *grads) {
    if (ln_grad(L->net, first, n, loss, 0)) return -1;
    return hipMemcpy(grads, L->net.g, LN_NPARAM * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

// one Adam step on items [first, first + n) of the data
int av1mi_ln_train_step(av1mi_learner *L, uint32_t first, uint32_t n, float lr, float wd, float *loss) {
    LnNet &N = L->net;
    if (ln_grad(N, first, n, loss, 0)) return -1;
    N.step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)N.step), c2 = 1.f - powf(b2, (float)N.step);
    hipLaunchKernelGGL(k_ln_adam, dim3((unsigned)((LN_NPARAM + 255) / 256)), dim3(256), 0, 0, N.w, N.g, N.m, N.v, (bf16_t *)nullptr,
                       (uint32_t)LN_NPARAM, lr, b1, b2, 1e-8f, c1, c2, wd);
    ln_refresh_bf16(N, 0);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}

}  // extern "C"
