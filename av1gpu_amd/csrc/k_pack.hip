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

}  // namespace av1m
