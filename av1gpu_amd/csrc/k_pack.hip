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
