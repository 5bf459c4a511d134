// Pack: per-SB output slots of the closed loop -> DecisionRecord arrays in decode order (tiles in raster order, SBs in
// raster order inside each tile), block_txb and coeff_off made image-relative.  One workgroup per SB; the offsets come
// from a host prefix sum over the SB counts.  Reference: encoder/src/k_misc.cu:46-106, pipeline.cpp:878-904.
#include <algorithm>
#include "av1m_common.h"

namespace av1m {

__global__ void __launch_bounds__(256) k_pack(Batch B, const PackOff *off, RecBlock *oB, uint32_t *oBT, RecTxb *oT, int16_t *oC) {
    const uint32_t sb = blockIdx.x;
    const SbOut so = B.sbOut[sb];
    const PackOff o = off[sb];
    for (uint32_t k = threadIdx.x; k < so.nBlocks; k += blockDim.x) {
        oB[o.gBlk + k] = B.oBlocks[(size_t)sb * MAX_LEAVES + k];
        oBT[o.gBlk + k] = o.rTxb + B.oBlockTxb[(size_t)sb * MAX_LEAVES + k];
    }
    for (uint32_t k = threadIdx.x; k < so.nTxb; k += blockDim.x) {
        RecTxb t = B.oTxb[(size_t)sb * B.txbCap + k];
        t.coeff_off += o.rCoef;
        oT[o.gTxb + k] = t;
    }
    const int16_t *src = B.oCoeffs + (size_t)sb * B.coefCap;
    for (uint32_t k = threadIdx.x; k < so.nCoeffs; k += blockDim.x) oC[o.gCoef + k] = src[k];
}

void launch_pack(const Batch &B, const PackOff *dOff, RecBlock *oB, uint32_t *oBT, RecTxb *oT, int16_t *oC, hipStream_t st) {
    if (!B.nSb) return;
    hipLaunchKernelGGL(k_pack, dim3(B.nSb), dim3(256), 0, st, B, dOff, oB, oBT, oT, oC);
}

// Sparse form of the packed coefficients, for the download to host memory (about 98 % of the levels are 0).  The
// array is cut into spans of SP_SPAN coefficients; lane l of a wave holds coefficients 8 l .. 8 l + 7 of a span, and
// mask word 8 s + e has bit l set when coefficient SP_SPAN s + 8 l + e is nonzero.  The nonzero values follow in the
// order (s, e, l).  A wave handles a run of SP_RUN spans: k_sparse_count counts each run's nonzeros, k_sparse_scan makes
// the runs' offsets (and the total), k_sparse_pack writes the masks and the values.
constexpr int SP_SPAN = 512, SP_RUN = 32;

__device__ __forceinline__ void sp_load(const int16_t *c, uint64_t n, uint64_t s, int lane, int (&v)[8]) {
    const uint64_t g = s * SP_SPAN + 8 * (uint64_t)lane;
    if (s * SP_SPAN + SP_SPAN <= n) {   // a whole span: one 16-byte load per lane (the array is 256-byte aligned)
        const uint4 w = *(const uint4 *)(c + g);
        const uint32_t u[4] = {w.x, w.y, w.z, w.w};
        for (int k = 0; k < 4; k++) { v[2 * k] = (int16_t)(u[k] & 0xFFFFu); v[2 * k + 1] = (int16_t)(u[k] >> 16); }
    } else {
        for (int e = 0; e < 8; e++) v[e] = g + e < n ? c[g + e] : 0;
    }
}

__global__ void __launch_bounds__(256) k_sparse_count(const int16_t *c, uint64_t n, uint64_t nSpan, uint32_t nRun, uint32_t *cnt) {
    const uint32_t run = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (run >= nRun) return;
    uint32_t k = 0;
    const uint64_t s1 = std::min<uint64_t>((uint64_t)run * SP_RUN + SP_RUN, nSpan);
    for (uint64_t s = (uint64_t)run * SP_RUN; s < s1; s++) {
        int v[8];
        sp_load(c, n, s, lane, v);
        for (int e = 0; e < 8; e++) k += (uint32_t)__popcll(__ballot(v[e] != 0));
    }
    if (lane == 0) cnt[run] = k;
}

// exclusive prefix sums of the runs' counts, and the total at off[nRun]; one workgroup of 1024 threads
__global__ void __launch_bounds__(1024) k_sparse_scan(const uint32_t *cnt, uint32_t nRun, uint64_t *off) {
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x, per = (nRun + 1023) / 1024, r0 = std::min(t * per, nRun), r1 = std::min(r0 + per, nRun);
    uint64_t sum = 0;
    for (uint32_t r = r0; r < r1; r++) sum += cnt[r];
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        uint64_t acc = 0;
        for (int k = 0; k < 1024; k++) { const uint64_t x = part[k]; part[k] = acc; acc += x; }
        off[nRun] = acc;
    }
    __syncthreads();
    uint64_t acc = part[t];
    for (uint32_t r = r0; r < r1; r++) { off[r] = acc; acc += cnt[r]; }
}

__global__ void __launch_bounds__(256) k_sparse_pack(const int16_t *c, uint64_t n, uint64_t nSpan, uint32_t nRun, const uint64_t *off,
                                                      uint64_t *mask, int16_t *val) {
    const uint32_t run = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (run >= nRun) return;
    uint64_t base = off[run];
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t s1 = std::min<uint64_t>((uint64_t)run * SP_RUN + SP_RUN, nSpan);
    for (uint64_t s = (uint64_t)run * SP_RUN; s < s1; s++) {
        int v[8];
        sp_load(c, n, s, lane, v);
        uint64_t mine = 0;
        for (int e = 0; e < 8; e++) {
            const uint64_t m = __ballot(v[e] != 0);
            if (v[e]) val[base + (uint64_t)__popcll(m & below)] = (int16_t)v[e];
            base += (uint64_t)__popcll(m);
            if (lane == e) mine = m;
        }
        if (lane < 8) mask[s * 8 + lane] = mine;
    }
}

// n coefficients at c (256-byte aligned) -> mask (ceil(n / SP_SPAN) * 8 words), cnt (nRun words), off (nRun + 1
// words: the runs' value offsets and the total), val (at most n values)
void launch_sparse(const int16_t *c, uint64_t n, uint64_t *mask, uint32_t *cnt, uint64_t *off, int16_t *val, hipStream_t st) {
    const uint64_t nSpan = (n + SP_SPAN - 1) / SP_SPAN;
    const uint32_t nRun = (uint32_t)((nSpan + SP_RUN - 1) / SP_RUN);
    if (!nRun) return;
    const dim3 grid((nRun + 3) / 4);
    hipLaunchKernelGGL(k_sparse_count, grid, dim3(256), 0, st, c, n, nSpan, nRun, cnt);
    hipLaunchKernelGGL(k_sparse_scan, dim3(1), dim3(1024), 0, st, cnt, nRun, off);
    hipLaunchKernelGGL(k_sparse_pack, grid, dim3(256), 0, st, c, n, nSpan, nRun, (const uint64_t *)off, mask, val);
}

// zero n 16-byte words (the encode arena's zeroed region: a kernel, so the step can live in a captured graph)
__global__ void __launch_bounds__(256) k_zero(uint4 *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = make_uint4(0u, 0u, 0u, 0u);
}
void launch_zero(void *p, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    if (!n) return;
    const size_t grid = std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)grid), dim3(256), 0, st, (uint4 *)p, n);
}

// 8-bit planes to 10-bit ones (the reference's gp_frames_to_10bit mapping, pipeline.h:218-225, libavif's full range):
// Y10 = round(Y8 * 1023 / 255), C10 = clamp(round((C8 - 128) * 1023 / 255) + 512, 0, 1023); no ties occur (1023 v / 255
// is never an odd multiple of 1/2).  Four samples per thread along a row, rows on the grid's y dimension.
__global__ void __launch_bounds__(256) k_to10(const uint8_t *src, int sp, uint16_t *dst, int dp, int w, int h, int chroma) {
    const int x0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    for (int y = blockIdx.y; y < h; y += gridDim.y) {
        const uint8_t *s = src + (size_t)y * sp;
        uint16_t *d = dst + (size_t)y * dp;
        for (int k = 0; k < 4; k++) {
            const int x = x0 + k;
            if (x >= w) break;
            const int v = s[x];
            int o;
            if (chroma) {
                const int dv = v - 128, m = ((dv < 0 ? -dv : dv) * 2046 + 255) / 510;
                o = 512 + (dv < 0 ? -m : m);
                o = o < 0 ? 0 : o > 1023 ? 1023 : o;
            } else {
                o = (v * 2046 + 255) / 510;
            }
            d[x] = (uint16_t)o;
        }
    }
}
void launch_to10(const uint8_t *src, int sp, uint16_t *dst, int dp, int w, int h, int chroma, hipStream_t st) {
    if (w <= 0 || h <= 0) return;
    hipLaunchKernelGGL(k_to10, dim3((unsigned)((w + 1023) / 1024), (unsigned)std::min(h, 4096)), dim3(256), 0, st, src, sp, dst, dp, w,
                       h, chroma);
}

}  // namespace av1m
