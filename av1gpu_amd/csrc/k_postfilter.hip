// Decoder-exact post-filters of an intra frame on the reconstruction (the reference's av1core/include/av1g/postfilter.h
// and av1core/src/gpu_postfilter.cu; AV1 spec 7.14 deblocking, 7.15 CDEF), for every image of a batch in one launch each.
//
// k_deblock: one thread per line of one 4-sample edge segment (pass 0: the vertical edges, consecutive threads on
//   consecutive segments of one row; pass 1: the horizontal edges, consecutive threads on consecutive columns, so the
//   column reads coalesce).  Inside a pass no two segments touch the same samples (a segment's filter reaches at most
//   its own transform block's half on either side), so the pass is one launch; pass 1 follows pass 0.
// k_cdef: one 256-thread workgroup per 64x64 superblock: the deblocked samples with a 2-sample border in LDS (-1 where
//   the tap lies outside the frame's 4x4 grid), the direction and variance of each 8x8 luma block by 64 threads, then
//   every sample of the three planes filtered into the separate output planes (unfiltered blocks copied).
#include "av1m_common.h"

namespace av1m {

// LoopfilterTxSizes (spec 7.14.1): luma the block's transform size, chroma the transform size of the block that carries
// the chroma of that 4x4 (the bottom-right luma 4x4 of an 8x8 in 4:2:0), both from the closed loop's 4x4 maps
__device__ __forceinline__ int lf_tx(const Img &I, int plane, int r4, int c4) {
    if (plane == 0) return I.mTx[(size_t)r4 * I.MiCols + c4];
    const int r = I.ss ? 2 * r4 + 1 : r4, c = I.ss ? 2 * c4 + 1 : c4;
    return uv_tx_size(I.mMiSize[(size_t)r * I.MiCols + c], I.ss);
}

// sample filtering process (7.14.6) of one line across an edge: q points at q0, step reaches q1 (and -step p0)
template <typename Px>
__device__ void lf_line(Px *q, ptrdiff_t step, int plane, int filterSize, int lvl, int sharp, int bd) {
    auto S = [&](int k) -> int { return (int)q[k * step]; };
    const int sh = bd - 8;
    const int shift = sharp > 4 ? 2 : sharp > 0 ? 1 : 0;
    const int lim0 = sharp > 0 ? clip3(1, 9 - sharp, lvl >> shift) : max(1, lvl >> shift);
    const int limit = lim0 << sh, blimit = (2 * (lvl + 2) + lim0) << sh, thresh = (lvl >> 4) << sh, flat = 1 << sh;
    const int len = filterSize == 4 ? 4 : plane ? 6 : filterSize == 8 ? 8 : 16;
    const int p0 = S(-1), p1 = S(-2), q0 = S(0), q1 = S(1);
    bool over = abs(p1 - p0) > limit || abs(q1 - q0) > limit || abs(p0 - q0) * 2 + abs(p1 - q1) / 2 > blimit;
    int p2 = 0, q2 = 0, p3 = 0, q3 = 0;
    if (len >= 6) { p2 = S(-3); q2 = S(2); over = over || abs(p2 - p1) > limit || abs(q2 - q1) > limit; }
    if (len >= 8) { p3 = S(-4); q3 = S(3); over = over || abs(p3 - p2) > limit || abs(q3 - q2) > limit; }
    if (over) return;   // filterMask 0
    const bool hev = abs(p1 - p0) > thresh || abs(q1 - q0) > thresh;
    bool isFlat = false, isFlat2 = false;
    if (filterSize >= 8) {
        isFlat = abs(p1 - p0) <= flat && abs(q1 - q0) <= flat && abs(p2 - p0) <= flat && abs(q2 - q0) <= flat;
        if (len >= 8) isFlat = isFlat && abs(p3 - p0) <= flat && abs(q3 - q0) <= flat;
    }
    if (filterSize >= 16 && isFlat)
        isFlat2 = abs(S(-7) - p0) <= flat && abs(S(6) - q0) <= flat && abs(S(-6) - p0) <= flat && abs(S(5) - q0) <= flat &&
                  abs(S(-5) - p0) <= flat && abs(S(4) - q0) <= flat;
    if (filterSize == 4 || !isFlat) {   // narrow filter (7.14.6.3)
        const int lo = -(1 << (bd - 1)), hi = (1 << (bd - 1)) - 1, off = 0x80 << sh;
        const int ps1 = p1 - off, ps0 = p0 - off, qs0 = q0 - off, qs1 = q1 - off;
        int f = hev ? clip3(lo, hi, ps1 - qs1) : 0;
        f = clip3(lo, hi, f + 3 * (qs0 - ps0));
        const int f1 = clip3(lo, hi, f + 4) >> 3, f2 = clip3(lo, hi, f + 3) >> 3;
        q[0] = (Px)(clip3(lo, hi, qs0 - f1) + off);
        q[-step] = (Px)(clip3(lo, hi, ps0 + f2) + off);
        if (!hev) {
            const int f3 = round2(f1, 1);
            q[step] = (Px)(clip3(lo, hi, qs1 - f3) + off);
            q[-2 * step] = (Px)(clip3(lo, hi, ps1 + f3) + off);
        }
        return;
    }
    // wide filter (7.14.6.4): n taps each side, the centre 2 n2 + 1 taps doubled, Round2 by log2Size
    const int log2Size = (filterSize == 8 || !isFlat2) ? 3 : 4;
    const int n = log2Size == 4 ? 6 : plane == 0 ? 3 : 2, n2 = (log2Size == 3 && plane == 0) ? 0 : 1;
    int v[14], out[12];
    for (int k = -(n + 1); k <= n; k++) v[k + 7] = S(k);
    for (int i = -n; i < n; i++) {
        int t = 0;
        for (int j = -n; j <= n; j++) t += v[clip3(-(n + 1), n, i + j) + 7] * (abs(j) <= n2 ? 2 : 1);
        out[i + 6] = round2(t, log2Size);
    }
    for (int i = -n; i < n; i++) q[i * step] = (Px)out[i + 6];
}

// Edge loop filter (7.14.2) of every image: pass 0 (vertical edges) or 1 (horizontal edges), one thread per line of a
// 4-sample segment; start[i] = first unit of image i (units of an image: its planes in order, pass 0 rows x segment
// columns, pass 1 segment rows x columns)
template <typename Px>
__global__ void __launch_bounds__(256) k_deblock(Batch B, int pass, const uint64_t *start) {
    const uint64_t u = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (u >= start[B.nimg]) return;
    int lo = 0, hi = B.nimg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= u) lo = mid; else hi = mid - 1;
    }
    const Img &I = B.img[lo];
    uint64_t k = u - start[lo];
    const int np = I.mono ? 1 : 3;
    for (int p = 0; p < np; p++) {
        const int ss = p ? I.ss : 0, rows = (I.MiRows * 4) >> ss, cols = (I.MiCols * 4) >> ss;
        const uint64_t n = pass == 0 ? (uint64_t)rows * (cols >> 2) : (uint64_t)(rows >> 2) * cols;
        if (k >= n) { k -= n; continue; }
        const int lvl = I.lfLevel[p == 0 ? pass : p + 1];
        if (!lvl) return;
        // the segment's 4x4 (plane units) and this thread's line
        int r4, c4, line;
        if (pass == 0) { line = (int)(k / (uint64_t)(cols >> 2)); c4 = (int)(k % (uint64_t)(cols >> 2)); r4 = line >> 2; }
        else { r4 = (int)(k / (uint64_t)cols); line = (int)(k % (uint64_t)cols); c4 = line >> 2; }
        const int xl = (c4 << ss) * 4, yl = (r4 << ss) * 4;   // luma position of the segment
        if (xl >= I.W || yl >= I.H) return;
        if ((pass == 0 && xl == 0) || (pass == 1 && yl == 0)) return;
        const int tx = lf_tx(I, p, r4, c4), prevTx = lf_tx(I, p, r4 - (pass == 1), c4 - (pass == 0));
        const int cur = pass == 0 ? av1t::Tx_Width[tx] : av1t::Tx_Height[tx];
        const int prev = pass == 0 ? av1t::Tx_Width[prevTx] : av1t::Tx_Height[prevTx];
        if (((pass == 0 ? c4 : r4) * 4) & (cur - 1)) return;   // not a transform edge (applyFilter: intra frames)
        const int filterSize = min(min(cur, prev), p ? 8 : 16);
        Px *pl = (Px *)I.rec[p];
        const int pitch = I.pitch[p];
        Px *q = pass == 0 ? pl + (size_t)line * pitch + c4 * 4 : pl + (size_t)(r4 * 4) * pitch + line;
        lf_line<Px>(q, pass == 0 ? 1 : pitch, p, filterSize, lvl, I.lfSharp, I.bd);
        return;
    }
}

// ---- CDEF (7.15) ----
constexpr int8_t Cdef_Dir[8][2][2] = {{{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}}, {{0, 1}, {1, 2}},
                                      {{1, 1}, {2, 2}},   {{1, 0}, {2, 1}},  {{1, 0}, {2, 0}}, {{1, 0}, {2, -1}}};
constexpr int Cdef_Div[9] = {0, 840, 420, 280, 210, 168, 140, 120, 105};

// constrain (7.15.3)
__device__ __forceinline__ int cdef_constrain(int diff, int threshold, int damping) {
    if (!threshold) return 0;
    const int adj = max(0, damping - (31 - __clz(threshold)));
    const int a = abs(diff), v = clip3(0, a, threshold - (a >> adj));
    return diff < 0 ? -v : v;
}

// the direction (7.15.2) of the 8x8 luma block at t (an LDS tile row pitch tp, samples >> (BitDepth - 8)) and its
// variance
__device__ __forceinline__ void cdef_direction(const int16_t *t, int tp, int bd, int &dir, int &var) {
    int cost[8];
    auto x = [&](int i, int j) { return (t[i * tp + j] >> (bd - 8)) - 128; };
    for (int d = 0; d < 8; d++) {
        int part[15];
        for (int k = 0; k < 15; k++) part[k] = 0;
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) {
                const int k = d == 0 ? i + j : d == 1 ? i + j / 2 : d == 2 ? i : d == 3 ? 3 + i - j / 2 : d == 4 ? 7 + i - j
                            : d == 5 ? 3 - i / 2 + j : d == 6 ? j : i / 2 + j;
                part[k] += x(i, j);
            }
        int c = 0;
        if (d == 2 || d == 6) {
            for (int i = 0; i < 8; i++) c += part[i] * part[i];
            c *= Cdef_Div[8];
        } else if (d == 0 || d == 4) {
            for (int i = 0; i < 7; i++) c += (part[i] * part[i] + part[14 - i] * part[14 - i]) * Cdef_Div[i + 1];
            c += part[7] * part[7] * Cdef_Div[8];
        } else {
            for (int j = 0; j < 5; j++) c += part[3 + j] * part[3 + j];
            c *= Cdef_Div[8];
            for (int j = 0; j < 3; j++) c += (part[j] * part[j] + part[10 - j] * part[10 - j]) * Cdef_Div[2 * j + 2];
        }
        cost[d] = c;
    }
    int best = 0, d = 0;
    for (int i = 0; i < 8; i++)
        if (cost[i] > best) { best = cost[i]; d = i; }
    dir = d;
    var = (best - cost[(d + 4) & 7]) >> 10;
}

// one sample of the CDEF filter (7.15.3) at tile position (y, x) of tile t (pitch tp; -1 = unavailable)
__device__ __forceinline__ int cdef_sample(const int16_t *t, int tp, int y, int x, int pri, int sec, int damping, int dir, int bd) {
    const int x0 = t[y * tp + x];
    const int odd = (pri >> (bd - 8)) & 1;
    int sum = 0, mx = x0, mn = x0;
    for (int k = 0; k < 2; k++)
        for (int s = -1; s <= 1; s += 2)
            for (int j = 0; j < 3; j++) {
                const int dd = j == 0 ? dir : (dir + (j == 1 ? -2 : 2)) & 7;
                const int v = t[(y + s * Cdef_Dir[dd][k][0]) * tp + x + s * Cdef_Dir[dd][k][1]];
                if (v < 0) continue;
                if (j == 0) sum += (odd ? 3 : k == 0 ? 4 : 2) * cdef_constrain(v - x0, pri, damping);
                else sum += (k == 0 ? 2 : 1) * cdef_constrain(v - x0, sec, damping);
                mx = max(mx, v);
                mn = min(mn, v);
            }
    return clip3(mn, mx, x0 + ((8 + sum - (sum < 0)) >> 4));
}

struct CdefShared {
    int16_t y[68 * 68];          // luma, 2-sample border
    int16_t c[2][68 * 68];       // chroma (36 x 36 in 4:2:0)
    int16_t bp[64][6];           // per 8x8: luma primary (-1: not filtered), secondary, direction; chroma the same
    int anyCoded;
};

// CDEF of superblock blockIdx.x: the deblocked planes (rec) into cdefOut; cdef_bits 0: one strength set, used by every
// superblock with a coded (non-skip) block, the others copied
template <typename Px>
__global__ void __launch_bounds__(256) k_cdef(Batch B) {
    __shared__ CdefShared S;
    const uint32_t sb = blockIdx.x;
    const Img &I = B.img[B.sbImg[sb]];
    if (!I.cdefOn) return;
    const int tid = threadIdx.x, local = (int)(sb - I.sbBase), sbc = local % I.sbCols, sbr = local / I.sbCols, bd = I.bd;
    const int np = I.mono ? 1 : 3;
    if (tid == 0) S.anyCoded = 0;
    __syncthreads();
    {
        const int r = sbr * 16 + (tid >> 4), c = sbc * 16 + (tid & 15);
        if (r < I.MiRows && c < I.MiCols && !I.mSkip[(size_t)r * I.MiCols + c]) S.anyCoded = 1;
    }
    __syncthreads();
    const bool on = S.anyCoded;
    // the tiles: plane sample (X, Y) -> entry (Y - y0 + 2) * tp + X - x0 + 2, -1 outside the 4x4 grid of the frame
    for (int p = 0; p < np && on; p++) {
        const int ss = p ? I.ss : 0, n = 64 >> ss, tp = n + 4, x0 = sbc * n, y0 = sbr * n;
        const int pw = (I.MiCols * 4) >> ss, ph = (I.MiRows * 4) >> ss;
        const Px *src = (const Px *)I.rec[p];
        int16_t *t = p == 0 ? S.y : S.c[p - 1];
        for (int i = tid; i < tp * tp; i += 256) {
            const int X = x0 - 2 + i % tp, Y = y0 - 2 + i / tp;
            t[i] = (X >= 0 && Y >= 0 && X < pw && Y < ph) ? (int16_t)src[(size_t)Y * I.pitch[p] + X] : (int16_t)-1;
        }
    }
    __syncthreads();
    if (on && tid < 64) {   // block parameters (7.15.1)
        const int r = sbr * 16 + 2 * (tid >> 3), c = sbc * 16 + 2 * (tid & 7);
        int16_t *o = S.bp[tid];
        o[0] = -1;
        if (r < I.MiRows && c < I.MiCols) {
            const uint8_t *sk = I.mSkip;
            const size_t m = (size_t)I.MiCols;
            const bool allSkip = sk[r * m + c] && sk[(r + 1) * m + c] && sk[r * m + c + 1] && sk[(r + 1) * m + c + 1];
            if (!allSkip) {
                int dir, var;
                cdef_direction(S.y + (2 + 8 * (tid >> 3)) * 68 + 2 + 8 * (tid & 7), 68, bd, dir, var);
                const int pri = I.cdefYPri << (bd - 8);
                int vs = 0;
                if (var >> 6) vs = min(31 - __clz(var >> 6), 12);
                const int uvPri = I.cdefUVPri << (bd - 8);
                o[0] = (int16_t)(var ? (pri * (4 + vs) + 8) >> 4 : 0);
                o[1] = (int16_t)(I.cdefYSec << (bd - 8));
                o[2] = (int16_t)(pri ? dir : 0);
                o[3] = (int16_t)uvPri;
                o[4] = (int16_t)(I.cdefUVSec << (bd - 8));
                o[5] = (int16_t)(uvPri ? dir : 0);
            }
        }
    }
    __syncthreads();
    for (int p = 0; p < np; p++) {
        const int ss = p ? I.ss : 0, n = 64 >> ss, tp = n + 4, x0 = sbc * n, y0 = sbr * n, bs = 8 >> ss;
        const int pw = (I.MiCols * 4) >> ss, ph = (I.MiRows * 4) >> ss;
        const Px *src = (const Px *)I.rec[p];
        Px *dst = (Px *)I.cdefOut[p];
        const int damping = I.cdefDamping + (bd - 8) - (p > 0);
        const int16_t *t = p == 0 ? S.y : S.c[p - 1];
        for (int i = tid; i < n * n; i += 256) {
            const int y = i / n, x = i % n;
            const size_t at = (size_t)(y0 + y) * I.pitch[p] + x0 + x;
            int v = src[at];
            if (on && x0 + x < pw && y0 + y < ph) {
                const int16_t *o = S.bp[(y / bs) * 8 + x / bs];
                if (o[0] >= 0)
                    v = p == 0 ? cdef_sample(t, tp, y + 2, x + 2, o[0], o[1], damping, o[2], bd)
                               : cdef_sample(t, tp, y + 2, x + 2, o[3], o[4], damping, o[5], bd);
            }
            dst[at] = (Px)v;
        }
    }
}

void launch_deblock(const Batch &B, int pass, const uint64_t *dStart, uint64_t units, int bd, hipStream_t st) {
    if (!units) return;
    const dim3 grid((unsigned)((units + 255) / 256));
    if (bd == 10) hipLaunchKernelGGL(k_deblock<uint16_t>, grid, dim3(256), 0, st, B, pass, dStart);
    else hipLaunchKernelGGL(k_deblock<uint8_t>, grid, dim3(256), 0, st, B, pass, dStart);
}
void launch_cdef(const Batch &B, int bd, hipStream_t st) {
    if (!B.nSb) return;
    if (bd == 10) hipLaunchKernelGGL(k_cdef<uint16_t>, dim3(B.nSb), dim3(256), 0, st, B);
    else hipLaunchKernelGGL(k_cdef<uint8_t>, dim3(B.nSb), dim3(256), 0, st, B);
}

}  // namespace av1m
