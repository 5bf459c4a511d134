// Host orchestration of one batch of frames on one MI355X (the reference's GpEncoder::init / encode,
// encoder/src/pipeline.cpp:221-1181, for the plain intra RD path) and the C ABI of include/av1mi.h.
//
//   encode: per-image descriptors and shape grids -> one device arena (stream-ordered) -> analysis (19 shape
//   launches) -> partition DP -> closed loop in SB wavefront steps (step = col + 2 row inside each tile) -> per-SB
//   counts to the host, prefix sums in decode order -> pack -> records and reconstruction to the host.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <sys/mman.h>
#include <algorithm>
#include <map>
#include <memory>
#include <atomic>
#include <thread>
#include <tuple>
#include <string>
#include <vector>
#include "../../include/av1mi.h"
#include "av1m_common.h"

namespace av1m {
void build_tables(DevTables &T);
void launch_analysis(const Batch &B, const SearchParams &P, const uint32_t *shTotal, int bd, const AnHooks &h, hipStream_t st);
void launch_analysis_on(const Batch &B, const SearchParams &P, const uint32_t *shTotal, int bd, const AnHooks &h, const hipStream_t *sts,
                        int ns);
void launch_zero(void *p, size_t bytes, hipStream_t st);
void launch_dp(const Batch &B, const SearchParams &P, hipStream_t st);
void launch_recon(const Batch &B, const SearchParams &P, const uint32_t *dSbList, uint32_t n, int bd, hipStream_t st);
void launch_recon_queue(const Batch &B, const SearchParams &P, const RcQueue &Q, uint32_t blocks, int bd, int wpe, hipStream_t st);
int recon_queue_slots(int bd, int wpe);
struct LnNet;
LnNet *ln_new(const float *params);
void ln_delete(LnNet *N);
int ln_masks(const LnNet &N, void *scratch, const Img *dImg, const uint32_t *dSbImg, uint32_t nSb, float conf, uint32_t chunk,
             uint32_t *dMask, hipStream_t st);
size_t ln_scratch_bytes(uint32_t n);
int recon_prepare(int bd);
size_t recon_smem_bytes();
void launch_pack(const Batch &B, const PackOff *dOff, RecBlock *oB, uint32_t *oBT, RecTxb *oT, int16_t *oC, hipStream_t st);
void launch_sparse(const int16_t *c, uint64_t n, uint64_t *mask, uint32_t *cnt, uint64_t *off, int16_t *val, hipStream_t st);
void launch_to10(const uint8_t *src, int sp, uint16_t *dst, int dp, int w, int h, int chroma, hipStream_t st);
hipError_t noalloc_scratch(void **ptr, size_t *bytes);
void launch_deblock(const Batch &B, int pass, const uint64_t *dStart, uint64_t units, int bd, hipStream_t st);
void launch_cdef(const Batch &B, int bd, hipStream_t st);
}  // namespace av1m

using namespace av1m;

static_assert(sizeof(av1mi_block) == 16 && sizeof(av1mi_txb) == 16, "DecisionRecord entries are 16 bytes");
static_assert(sizeof(RecBlock) == 16 && sizeof(RecTxb) == 16, "DecisionRecord entries are 16 bytes");
static_assert(sizeof(av1mi_params) == sizeof(SearchParams), "av1mi_params mirrors SearchParams");
static_assert(sizeof(av1mi_frame_params) == 88, "av1mi_frame_params has av1g_frame_params layout");

namespace {

#define HCK(x)                                                                                                 \
    do {                                                                                                       \
        const hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) {                                                                                \
            snprintf(err, errlen, "HIP: %s at %s (%s:%d)", hipGetErrorString(e_), #x, __FILE__, __LINE__);       \
            return -1;                                                                                         \
        }                                                                                                      \
    } while (0)

int tile_log2(int blk, int target) { int k = 0; while ((blk << k) < target) k++; return k; }
// minimum legal uniform tiling (spec tile_info; reference av1core/src/encoder.cpp:56-91)
void tile_layout(int W, int H, int &cols, int &rows, int *cs, int *rs) {
    const int MiCols = 2 * ((W + 7) >> 3), MiRows = 2 * ((H + 7) >> 3);
    const int sbCols = (MiCols + 15) >> 4, sbRows = (MiRows + 15) >> 4;
    const int maxW = 4096 >> 6, maxA = (4096 * 2304) >> 12;
    const int minC = tile_log2(maxW, sbCols);
    int minT = std::max(tile_log2(maxA, sbRows * sbCols), minC);
    const int tws = (sbCols + (1 << minC) - 1) >> minC;
    cols = 0;
    for (int s = 0; s < sbCols; s += tws) cs[cols++] = s << 4;
    cs[cols] = MiCols;
    for (int rl = std::max(minT - minC, 0);; rl++) {
        const int ths = (sbRows + (1 << rl) - 1) >> rl;
        if (tws * ths <= maxA || ths == 1) {
            rows = 0;
            for (int s = 0; s < sbRows; s += ths) rs[rows++] = s << 4;
            rs[rows] = MiRows;
            return;
        }
    }
}

// stream-ordered device arena of one encode call
struct Arena {
    std::vector<std::pair<size_t, void **>> req;
    size_t total = 0;
    void *base = nullptr;
    template <class T> void add(T *&p, size_t n) {
        req.push_back({total, (void **)&p});
        total += (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
    }
    bool owned = false;
    hipError_t alloc(hipStream_t st) {
        hipError_t e = hipMallocAsync(&base, total, st);
        if (e != hipSuccess) return e;
        owned = true;
        for (auto &r : req) *r.second = (char *)base + r.first;
        return hipSuccess;
    }
    void bind(void *mem) {   // lay the requests out over memory the caller owns (a captured plan's)
        base = mem;
        owned = false;
        for (auto &r : req) *r.second = (char *)base + r.first;
    }
    void release(hipStream_t st) {
        if (base && owned) (void)hipFreeAsync(base, st);
        base = nullptr;
        owned = false;
    }
};

}  // namespace

struct av1mi_encoder {
    av1mi_config cfg;
    int device = 0;
    DevTables *dT = nullptr;
    hipEvent_t ev[10] = {};
    std::vector<hipEvent_t> dlEv;     // ends of the segments of the record download (host copies start behind them)
    int cus = 0, slots[2][2] = {{0, 0}, {0, 0}};   // compute units; resident closed-loop workgroups per CU [10-bit][variant]
    uint64_t clockHz = 100000000;     // wall_clock64 rate
    LnNet *ln = nullptr;              // learned pruning network (null: exhaustive search)
    float lnConf = 0.8f;
    uint32_t lnChunk = 65536;         // superblocks per network pass (bounds the activation memory, ~100 KB each)
    void *stage = nullptr;            // pinned host staging of the records (grow-only, by doubling)
    size_t stageBytes = 0;
    std::vector<void *> oldStages;    // outgrown staging buffers, freed at destroy: hipHostFree would wait for the whole
                                      // device, i.e. for the other encoders' work in a multi-stream server
    uint64_t *spTot = nullptr;        // pinned: the nonzero count of a call's sparse coefficient download
    hipEvent_t spEv = nullptr;        // after that count's copy
    bool countWork = false;           // count the forward-transform FLOPs of the evaluations (av1mi_times.flops*)
    // analysis hooks (device pointers of the caller): external candidates per shape with their per-(shape, image) offsets,
    // block lists per shape
    const ExtCand *ext[NSHAPES] = {};
    const int64_t *extOff = nullptr;
    int extTopK = 0;
    const uint32_t *list[NSHAPES] = {}, *listCount[NSHAPES] = {};
    bool anyList = false;
    bool dpExport = false;            // keep the DP's per-node relative costs of the last call (training labels)
    std::vector<float> dpRel;
    void *scratch = nullptr;          // this device's copy of the no-allocation scratch array, and its size
    size_t scratchBytes = 0;
    bool searchExport = false;        // keep the last call's analysis records, superblock costs and leaves (checks)
    std::vector<Img> exImg;
    std::vector<BlockRes> exBlk;
    std::vector<float> exSbJ;
    std::vector<uint16_t> exLeafCount;
    std::vector<Leaf> exLeaves;
    hipStream_t aux[3] = {};          // analysis streams besides the caller's
    hipStream_t cap = nullptr;        // graph capture stream
    hipEvent_t evFork = nullptr, evJoin[3] = {};
    uint64_t gen = 0;                 // bumped when the captured state (network, export) changes
    struct Plan {                     // captured device part of one batch shape, over its own persistent arena
        std::vector<uint64_t> key;
        void *mem = nullptr;
        hipGraphExec_t exec = nullptr;
    } plan;
    // autotuner state of one batch shape: candidates tried, their device times.  Candidate k: analysis streams 1 (k
    // even) or 4; closed loop k >> 1 = 0 persistent queue at one workgroup per CU, 1 at two, 2 one launch per step
    static constexpr int NTUNE = 6;
    struct Tune {
        int next = 0, best = 0;
        float ms[NTUNE] = {};
    };
    std::map<std::vector<uint64_t>, Tune> tunes;
};

static void drop_plan(av1mi_encoder *e) {
    if (e->plan.exec) (void)hipGraphExecDestroy(e->plan.exec);
    if (e->plan.mem) (void)hipFree(e->plan.mem);
    e->plan = av1mi_encoder::Plan();
}

// a fresh output buffer of at least 4 MiB asks for transparent huge pages (where the kernel offers them on request), so
// its first touch by the copy faults once per 2 MiB instead of once per 4 KiB; zero: through calloc, whose large blocks
// are fresh zero pages (nothing written before the first touch)
static void *out_alloc(size_t bytes, bool zero = false) {
    void *p = zero ? calloc(bytes, 1) : malloc(bytes);
    constexpr size_t HP = (size_t)2 << 20;
    if (p && bytes >= 2 * HP) {
        const uintptr_t a = ((uintptr_t)p + HP - 1) & ~(uintptr_t)(HP - 1), e = ((uintptr_t)p + bytes) & ~(uintptr_t)(HP - 1);
        if (e > a) (void)madvise((void *)a, e - a, MADV_HUGEPAGE);
    }
    return p;
}

// Copy jobs (dst, src, bytes) out of the pinned staging buffer on up to 16 host threads, in chunks (which spreads the
// first-touch faults of fresh outputs), while the download into it is still running: the download is issued as
// segments in staging order, segEnd[k] the end offset of segment k and segEv[k] the event after it; the chunks are
// taken in staging order and each waits only for the segment that covers its end.
static int parallel_copy(const std::vector<std::tuple<void *, const void *, size_t>> &jobs, const char *stage,
                         const std::vector<size_t> &segEnd, const std::vector<hipEvent_t> &segEv) {
    constexpr size_t CH = (size_t)8 << 20;
    struct Part { char *d; const char *s; size_t n; size_t seg; };
    std::vector<Part> parts;
    size_t total = 0;
    for (const auto &j : jobs) {
        char *d = (char *)std::get<0>(j);
        const char *src = (const char *)std::get<1>(j);
        for (size_t o = 0; o < std::get<2>(j); o += CH) parts.push_back(Part{d + o, src + o, std::min(CH, std::get<2>(j) - o), 0});
        total += std::get<2>(j);
    }
    for (auto &pt : parts) {
        const size_t end = (size_t)(pt.s - stage) + pt.n;
        pt.seg = (size_t)(std::lower_bound(segEnd.begin(), segEnd.end(), end) - segEnd.begin());
        if (pt.seg >= segEnd.size()) return -1;
    }
    std::sort(parts.begin(), parts.end(), [](const Part &a, const Part &b) { return a.s < b.s; });
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = (unsigned)std::min<size_t>(std::min(16u, hw), std::max<size_t>(1, total / CH));
    std::atomic<size_t> next{0};
    std::atomic<int> bad{0};
    auto work = [&]() {
        size_t waited = 0;   // segments [0, waited) are known complete on this thread
        for (size_t k; (k = next.fetch_add(1)) < parts.size();) {
            const Part &pt = parts[k];
            while (waited <= pt.seg) {
                if (hipEventSynchronize(segEv[waited]) != hipSuccess) { bad = 1; return; }
                waited++;
            }
            memcpy(pt.d, pt.s, pt.n);
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    return bad ? -1 : 0;
}

// The frames' host coefficient arrays, made while the device works: allocated zeroed at an upper bound of the count
// (every superblock's area in each plane: a tx block stores at most its area and lies in its superblock) and faulted
// in on up to 8 host threads, so that after the download only the nonzero values remain to be written.  take() hands a
// frame's array over, shrunk to its count; what is not taken is freed with the object.
struct CoefPrefault {
    std::vector<int16_t *> buf;
    std::vector<uint64_t> cap;
    std::vector<std::thread> th;
    void start(const std::vector<Img> &imgs) {
        const size_t n = imgs.size();
        buf.assign(n, nullptr);
        cap.resize(n);
        std::vector<size_t> order(n);
        for (size_t i = 0; i < n; i++) {
            const Img &I = imgs[i];
            const uint64_t a = (uint64_t)I.sbCols * 64 * I.sbRows * 64;
            cap[i] = a + (I.mono ? 0 : 2 * (a >> (2 * I.ss)));
            order[i] = i;
        }
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cap[a] > cap[b]; });
        auto next = std::make_shared<std::atomic<size_t>>(0);
        const unsigned nt = (unsigned)std::min<size_t>(std::min(8u, std::max(1u, std::thread::hardware_concurrency())), n);
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([this, order, next]() {
                for (size_t k; (k = next->fetch_add(1)) < order.size();) {
                    const size_t i = order[k], bytes = cap[i] * 2;
                    volatile char *p = (volatile char *)out_alloc(bytes, true);
                    if (!p) continue;
                    for (size_t o = 0; o < bytes; o += 4096) p[o] = 0;
                    buf[i] = (int16_t *)p;
                }
            });
    }
    void join() {
        for (auto &t : th) t.join();
        th.clear();
    }
    int16_t *take(size_t i, uint64_t count) {
        join();
        if (i >= buf.size() || !buf[i] || count > cap[i]) return nullptr;
        int16_t *p = (int16_t *)realloc(buf[i], std::max<uint64_t>(count, 1) * 2);
        if (p) buf[i] = nullptr;
        return p;
    }
    ~CoefPrefault() {
        join();
        for (int16_t *p : buf) free(p);
    }
};

// The frames' coefficients from the sparse download (k_sparse_pack's layout: mask word 8 s + e has bit l set for a
// nonzero coefficient 512 s + 8 l + e of the batch's packed array; the values follow in the order (s, e, l), the run
// of 32 spans r from off[r]).  Frame i owns the packed coefficients [c0[i], c0[i + 1]) and its array dst[i] is zeroed,
// so only the nonzero values are written.  Groups of 8 runs go to up to 16 host threads; -1 if a group's values do not
// end at the next group's offset.
static int sparse_expand(const uint64_t *mask, const uint64_t *off, const int16_t *val, uint64_t n, const std::vector<uint64_t> &c0,
                         const std::vector<int16_t *> &dst) {
    constexpr uint64_t SPAN = 512, RUN = 32, GROUP = 8;
    const uint64_t nSpan = (n + SPAN - 1) / SPAN, nRun = (nSpan + RUN - 1) / RUN, nJob = (nRun + GROUP - 1) / GROUP;
    if (!nJob) return 0;
    std::atomic<uint64_t> next{0};
    std::atomic<int> bad{0};
    const auto frameOf = [&](uint64_t g) { return (size_t)(std::upper_bound(c0.begin(), c0.end(), g) - c0.begin()) - 1; };
    auto work = [&]() {
        for (uint64_t j; (j = next.fetch_add(1)) < nJob;) {
            const uint64_t r0 = j * GROUP, r1 = std::min(r0 + GROUP, nRun), s1 = std::min(r1 * RUN, nSpan);
            uint64_t v = off[r0];
            size_t f = frameOf(r0 * RUN * SPAN);
            for (uint64_t s = r0 * RUN; s < s1; s++) {
                const uint64_t lo = s * SPAN;
                while (c0[f + 1] <= lo) f++;
                const bool whole = lo + SPAN <= c0[f + 1];   // the span lies in frame f
                for (int e = 0; e < 8; e++)
                    for (uint64_t m = mask[s * 8 + e]; m; m &= m - 1) {
                        const uint64_t g = lo + 8 * (uint64_t)__builtin_ctzll(m) + (uint64_t)e;
                        const size_t h = whole ? f : frameOf(g);
                        dst[h][g - c0[h]] = val[v++];
                    }
            }
            if (v != off[r1]) bad = 1;
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = (unsigned)std::min<uint64_t>(std::min(16u, hw), nJob);
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    return bad ? -1 : 0;
}

extern "C" {

void av1mi_default_params(av1mi_params *p) {
    memset(p, 0, sizeof *p);
    p->lambdaMul = 1.f; p->satdLambda = 1.f;
    p->rnd = 0.375f; p->zbin = 0.656f; p->rndC = 0.375f; p->zbinC = 0.656f;
    p->K = 6; p->Krecheck = 6; p->txSearch = 1; p->txSearchCL = 1; p->txAllCL = 1; p->deltas = 1; p->uvDeltas = 1;
    p->cfl = 1; p->uvK = 6; p->shapeMask = 0x3F1FFF; p->partMask8 = 0xF; p->partMaskN = 0x3FF; p->maxPart = 64;
    p->uvOverheadBits = 1.5f; p->splitBias = 1.f; p->lambdaCMul = 1.f;
}

// Whether an n x n node at (x, y) of a superblock whose frame part is FW x FH can be partitioned into blocks the analysis
// evaluates, under the partitions the DP may choose there (the same edge rules and options as k_dp): recursively over
// Part_Shape, a partition being usable when each of its in-frame parts is an evaluated shape or a partitionable node.
static bool partitionable(const av1mi_params &P, int n, int x, int y, int FW, int FH) {
    if (x >= FW || y >= FH) return true;
    const int q = n >> 2;
    const bool rows = y + 2 * q < FH, cols = x + 2 * q < FW, big = n > P.maxPart;
    uint32_t allowed;
    if (rows && cols) allowed = big ? (1u << PARTITION_SPLIT) : ((uint32_t)(n == 8 ? P.partMask8 : P.partMaskN) | (1u << PARTITION_SPLIT));
    else if (cols) allowed = (1u << PARTITION_SPLIT) | (big ? 0u : 1u << PARTITION_HORZ);
    else if (rows) allowed = (1u << PARTITION_SPLIT) | (big ? 0u : 1u << PARTITION_VERT);
    else allowed = 1u << PARTITION_SPLIT;
    if (n == 8) allowed &= 0xFu;
    auto evaluated = [&](int bs) { return bs < BLOCK_INVALID && !(bs >= 13 && bs <= 15) && ((P.shapeMask >> bs) & 1); };
    for (int p = 0; p < 10; p++) {
        if (!((allowed >> p) & 1)) continue;
        bool ok = true;
        for (int k = 0; k < Part_Shape[p].parts && ok; k++) {
            const int px = x + Part_Shape[p].at[k].dx * q, py = y + Part_Shape[p].at[k].dy * q;
            if (px >= FW || py >= FH) continue;
            ok = (p == PARTITION_SPLIT && n > 8) ? partitionable(P, n >> 1, px, py, FW, FH) : evaluated(part_bsize(p, k, n));
        }
        if (ok) return true;
    }
    return false;
}

int av1mi_check_partitions(const av1mi_params *p, char *err, int errlen) {
    // every superblock, whatever part of it the frame covers (8 .. 64 samples each way), keeps a partition of finite
    // cost (gp_check_partitions, pipeline.cpp:136-219)
    for (int FH = 8; FH <= 64; FH += 8)
        for (int FW = 8; FW <= 64; FW += 8)
            if (!partitionable(*p, 64, 0, 0, FW, FH)) {
                snprintf(err, errlen, "the shape / partition options leave some superblock no partition the analysis evaluates "
                                      "(a frame part of %dx%d samples)", FW, FH);
                return -1;
            }
    return 0;
}

int av1mi_create(const av1mi_config *cfg, av1mi_encoder **out, char *err, int errlen) {
    *out = nullptr;
    av1mi_config c = *cfg;
    c.P.K = std::max(1, std::min(c.P.K, MAXK));
    c.P.Krecheck = std::max(1, std::min(c.P.Krecheck, c.P.K));
    if (c.qindex < 1 || c.qindex > 255) { snprintf(err, errlen, "qindex %d outside 1..255", c.qindex); return -1; }
    if (av1mi_check_partitions(&c.P, err, errlen)) return -1;
    av1mi_encoder *e = new av1mi_encoder();
    e->cfg = c;
    hipError_t he = hipGetDevice(&e->device);
    if (he != hipSuccess) { snprintf(err, errlen, "HIP: %s (no usable GPU)", hipGetErrorString(he)); delete e; return -1; }
    DevTables *h = (DevTables *)malloc(sizeof(DevTables));
    build_tables(*h);
    he = hipMalloc(&e->dT, sizeof(DevTables));
    if (he == hipSuccess) he = hipMemcpy(e->dT, h, sizeof(DevTables), hipMemcpyHostToDevice);
    free(h);
    if (he != hipSuccess) { snprintf(err, errlen, "HIP: %s (table upload)", hipGetErrorString(he)); delete e; return -1; }
    if (recon_prepare(8) || recon_prepare(10)) { snprintf(err, errlen, "closed-loop kernel needs %zu bytes of LDS", recon_smem_bytes()); delete e; return -1; }
    if (c.schedule < 0 || c.schedule > 1 || c.queueBlocks < 0 || c.queueOcc < 0 || c.queueOcc > 2) { snprintf(err, errlen, "bad schedule %d / queueBlocks %d", c.schedule, c.queueBlocks); delete e; return -1; }
    const auto secOk = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4; };
    if (c.lfMode < 0 || c.lfMode > 2 || (c.lfMode == 2 && (c.lfLevel < 0 || c.lfLevel > 63)) || !(c.lfMul >= 0.f && c.lfMul < 1e6f) ||
        c.cdef < 0 || c.cdef > 1 ||
        (c.cdef && (c.cdefDampingMinus3 < 0 || c.cdefDampingMinus3 > 3 || c.cdefYPri < 0 || c.cdefYPri > 15 || c.cdefUVPri < 0 ||
                    c.cdefUVPri > 15 || !secOk(c.cdefYSec) || !secOk(c.cdefUVSec)))) {
        snprintf(err, errlen, "bad post-filter options: lfMode %d (0..2), lfLevel %d (0..63), lfMul %g, cdef %d, damping-3 %d (0..3), "
                 "strengths Y %d/%d UV %d/%d (primary 0..15, secondary 0, 1, 2, 4)", c.lfMode, c.lfLevel, c.lfMul, c.cdef,
                 c.cdefDampingMinus3, c.cdefYPri, c.cdefYSec, c.cdefUVPri, c.cdefUVSec);
        delete e;
        return -1;
    }
    if (c.anStreams < 0 || c.anStreams > 4 || c.graph < 0 || c.graph > 1 || c.tune < 0 || c.tune > 1) {
        snprintf(err, errlen, "bad anStreams %d / graph %d / tune %d", c.anStreams, c.graph, c.tune);
        delete e;
        return -1;
    }
    // the no-allocation entry point's scratch array: this device's address, resolved with the device current
    if ((he = noalloc_scratch(&e->scratch, &e->scratchBytes)) != hipSuccess) {
        snprintf(err, errlen, "HIP: %s (scratch array)", hipGetErrorString(he));
        av1mi_destroy(e);
        return -1;
    }
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, e->device) == hipSuccess) e->cus = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeWallClockRate, e->device) == hipSuccess && v > 0) e->clockHz = (uint64_t)v * 1000;
    for (int b = 0; b < 2; b++)
        for (int v = 0; v < 2; v++) e->slots[b][v] = recon_queue_slots(b ? 10 : 8, v + 1);
    for (auto &v : e->ev)
        if ((he = hipEventCreate(&v)) != hipSuccess) { snprintf(err, errlen, "HIP: %s (event)", hipGetErrorString(he)); av1mi_destroy(e); return -1; }
    // default-priority non-blocking streams for the analysis branches and the capture; timing-free events joining them
    for (auto &s : e->aux)
        if ((he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) { snprintf(err, errlen, "HIP: %s (stream)", hipGetErrorString(he)); av1mi_destroy(e); return -1; }
    if ((he = hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking)) != hipSuccess) { snprintf(err, errlen, "HIP: %s (stream)", hipGetErrorString(he)); av1mi_destroy(e); return -1; }
    if ((he = hipEventCreateWithFlags(&e->evFork, hipEventDisableTiming)) != hipSuccess) { snprintf(err, errlen, "HIP: %s (event)", hipGetErrorString(he)); av1mi_destroy(e); return -1; }
    for (auto &v : e->evJoin)
        if ((he = hipEventCreateWithFlags(&v, hipEventDisableTiming)) != hipSuccess) { snprintf(err, errlen, "HIP: %s (event)", hipGetErrorString(he)); av1mi_destroy(e); return -1; }
    // keep freed stream-ordered allocations cached between batches
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, e->device) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    *out = e;
    return 0;
}

int av1mi_set_shortlist(av1mi_encoder *e, const float *params, float confidence, char *err, int errlen) {
    drop_plan(e);   // a captured graph holds the old network's pointers
    e->gen++;
    ln_delete(e->ln);
    e->ln = nullptr;
    if (!params) return 0;
    if (!(confidence > 0.5f && confidence <= 1.0f)) { snprintf(err, errlen, "confidence %g outside (0.5, 1]", confidence); return -1; }
    if (e->anyList) { snprintf(err, errlen, "the learned pruning and block lists do not combine"); return -1; }
    e->ln = ln_new(params);
    if (!e->ln) { snprintf(err, errlen, "HIP: could not upload the pruning network"); return -1; }
    e->lnConf = confidence;
    const char *c = getenv("AV1MI_LN_CHUNK");
    e->lnChunk = c && atoi(c) > 0 ? (uint32_t)atoi(c) : 65536u;
    return 0;
}

int av1mi_set_ext_candidates(av1mi_encoder *e, const av1mi_ext_cand *const *cand, const int64_t *off, int topK, char *err,
                             int errlen) {
    static_assert(sizeof(av1mi_ext_cand) == sizeof(ExtCand) && sizeof(ExtCand) == 8, "GpExtCand is 8 bytes");
    bool any = false;
    for (int s = 0; cand && s < NSHAPES; s++) any = any || cand[s];
    if (any && (!off || topK < 1 || topK > 32)) {
        snprintf(err, errlen, "external candidates need the offsets and 1..32 candidates per block (topK %d)", topK);
        return -1;
    }
    drop_plan(e);
    e->gen++;
    for (int s = 0; s < NSHAPES; s++) e->ext[s] = any ? (const ExtCand *)cand[s] : nullptr;
    e->extOff = any ? off : nullptr;
    e->extTopK = any ? topK : 0;
    return 0;
}

int av1mi_set_block_lists(av1mi_encoder *e, const uint32_t *const *list, const uint32_t *const *count, char *err, int errlen) {
    av1mi_params P = e->cfg.P;
    bool any = false;
    for (int s = 0; list && s < NSHAPES; s++)
        if (list[s]) {
            if (!count || !count[s]) { snprintf(err, errlen, "block list of shape %d without its count", s); return -1; }
            P.shapeMask &= ~(1 << s);
            any = true;
        }
    // every superblock must keep a partition of finite cost whatever the lists leave out
    if (any && av1mi_check_partitions(&P, err, errlen)) {
        const std::string m = std::string("block lists: the shapes without a list must leave every superblock a partition (") + err + ")";
        snprintf(err, errlen, "%s", m.c_str());
        return -1;
    }
    if (any && e->ln) { snprintf(err, errlen, "block lists and the learned pruning do not combine"); return -1; }
    drop_plan(e);
    e->gen++;
    for (int s = 0; s < NSHAPES; s++) { e->list[s] = any ? list[s] : nullptr; e->listCount[s] = any && list[s] ? count[s] : nullptr; }
    e->anyList = any;
    return 0;
}

int av1mi_set_work_count(av1mi_encoder *e, int on) {
    drop_plan(e);
    e->gen++;
    e->countWork = on != 0;
    return 0;
}

int av1mi_set_dp_export(av1mi_encoder *e, int on) {
    drop_plan(e);
    e->gen++;
    e->dpExport = on != 0;
    e->dpRel.clear();
    return 0;
}
size_t av1mi_dp_export(const av1mi_encoder *e, float *out, size_t n) {
    const size_t k = std::min(n, e->dpRel.size());
    if (out && k) memcpy(out, e->dpRel.data(), k * 4);
    return e->dpRel.size();
}

int av1mi_set_search_export(av1mi_encoder *e, int on) {
    drop_plan(e);
    e->gen++;
    e->searchExport = on != 0;
    e->exImg.clear(); e->exBlk.clear(); e->exSbJ.clear(); e->exLeafCount.clear(); e->exLeaves.clear();
    return 0;
}
size_t av1mi_blockres_export(const av1mi_encoder *e, int img, int shape, void *out, size_t n) {
    if (img < 0 || img >= (int)e->exImg.size() || shape < 0 || shape >= NSHAPES) return 0;
    const Img &I = e->exImg[img];
    const size_t cnt = (size_t)I.shNbx[shape] * I.shNby[shape];
    if (out) memcpy(out, e->exBlk.data() + I.shOff[shape], std::min(n, cnt) * sizeof(BlockRes));
    return cnt;
}
size_t av1mi_sbj_export(const av1mi_encoder *e, int img, float *out, size_t n) {
    if (img < 0 || img >= (int)e->exImg.size()) return 0;
    const Img &I = e->exImg[img];
    const size_t cnt = (size_t)I.sbCols * I.sbRows;
    if (out) memcpy(out, e->exSbJ.data() + I.sbBase, std::min(n, cnt) * 4);
    return cnt;
}
size_t av1mi_leaves_export(const av1mi_encoder *e, int img, uint16_t *counts, void *leaves, size_t nsb) {
    if (img < 0 || img >= (int)e->exImg.size()) return 0;
    const Img &I = e->exImg[img];
    const size_t cnt = (size_t)I.sbCols * I.sbRows, k = std::min(nsb, cnt);
    if (counts) memcpy(counts, e->exLeafCount.data() + I.sbBase, k * 2);
    if (leaves) memcpy(leaves, e->exLeaves.data() + (size_t)I.sbBase * MAX_LEAVES, k * MAX_LEAVES * sizeof(Leaf));
    return cnt;
}

void av1mi_destroy(av1mi_encoder *e) {
    if (!e) return;
    drop_plan(e);
    ln_delete(e->ln);
    if (e->stage) (void)hipHostFree(e->stage);
    for (void *p : e->oldStages) (void)hipHostFree(p);
    if (e->spTot) (void)hipHostFree(e->spTot);
    if (e->spEv) (void)hipEventDestroy(e->spEv);
    for (auto &s : e->aux) if (s) (void)hipStreamDestroy(s);
    if (e->cap) (void)hipStreamDestroy(e->cap);
    for (auto &v : e->dlEv) (void)hipEventDestroy(v);
    if (e->evFork) (void)hipEventDestroy(e->evFork);
    for (auto &v : e->evJoin) if (v) (void)hipEventDestroy(v);
    if (e->dT) (void)hipFree(e->dT);
    for (auto &v : e->ev) if (v) (void)hipEventDestroy(v);
    delete e;
}

void av1mi_get_config(const av1mi_encoder *e, av1mi_config *cfg) { *cfg = e->cfg; }

}  // extern "C"

// The post-filter decisions of a frame (pipeline.cpp:960-986): the loop filter level of lfMode 1 is libaom's
// LPF_PICK_FROM_Q fit for a key frame at the frame's ac quantizer (8-bit: Round2(acq * 17563 - 421574, 18); 10-bit:
// Round2(acq * 20723 + 4060632, 20) - 4) times lfMul, rounded to nearest (ties to even) and clamped to 0..63; the same
// level on every plane (none on chroma of monochrome frames), sharpness 0; CDEF one strength set.
static void frame_postfilter(const av1mi_config &c, Img &I) {
    memset(I.lfLevel, 0, sizeof I.lfLevel);
    I.lfSharp = 0;
    if (c.lfMode) {
        int lv = c.lfLevel;
        const long acq = ac_q(I.qindex, I.bd);
        if (c.lfMode == 1) {
            const long v = I.bd == 10 ? ((acq * 20723 + 4060632 + (1 << 19)) >> 20) - 4 : (acq * 17563 - 421574 + (1 << 17)) >> 18;
            lv = (int)lrintf((float)v * c.lfMul);
        }
        lv = std::max(0, std::min(63, lv));
        I.lfLevel[0] = I.lfLevel[1] = (uint8_t)lv;
        I.lfLevel[2] = I.lfLevel[3] = (uint8_t)(I.mono ? 0 : lv);
    }
    I.cdefOn = (uint8_t)(c.cdef ? 1 : 0);
    I.cdefDamping = (uint8_t)(c.cdefDampingMinus3 + 3);
    I.cdefYPri = (uint8_t)c.cdefYPri; I.cdefYSec = (uint8_t)c.cdefYSec;
    I.cdefUVPri = (uint8_t)c.cdefUVPri; I.cdefUVSec = (uint8_t)c.cdefUVSec;
}

// out.fp: the frame parameters of av1mi_frame_params_of with the frame's post-filter decisions
static void frame_params_fill(av1mi_frame_out &o, const Img &I) {
    av1mi_frame_params_of(&o, I.edgeFilter, -1, &o.fp);
    for (int k = 0; k < 4; k++) o.fp.loop_filter_level[k] = I.lfLevel[k];
    o.fp.loop_filter_sharpness = I.lfSharp;
    o.fp.enable_cdef = I.cdefOn;
    if (I.cdefOn) {
        o.fp.cdef_damping_minus_3 = (uint8_t)(I.cdefDamping - 3);
        o.fp.cdef_bits = 0;
        o.fp.cdef_y_pri[0] = I.cdefYPri; o.fp.cdef_y_sec[0] = I.cdefYSec;
        o.fp.cdef_uv_pri[0] = I.cdefUVPri; o.fp.cdef_uv_sec[0] = I.cdefUVSec;
    }
}

// The no-allocation variant of an encode: the arena laid over the device's scratch array, the packed records written
// into the caller's buffers.
struct NoAlloc {
    void *mem;
    size_t bytes;
    const av1mi_noalloc_out *o;
};

// GpEncoder::encode (pipeline.cpp:323-1181) for the plain intra path.  dev != nullptr: the packed records stay on the
// device and pass to the caller (GpDevRecords); na != nullptr: no device allocation (scratch array, caller's outputs);
// otherwise they are copied into each frame's host arrays.
static int encode_impl(av1mi_encoder *enc, const av1mi_frame_in *in, int n, av1mi_frame_out *out, av1mi_times *t, void *stream,
                       int wantRecon, av1mi_dev_records *dev, const NoAlloc *na, char *err, int errlen) {
    hipStream_t st = (hipStream_t)stream;
    av1mi_times tm;
    memset(&tm, 0, sizeof tm);
    for (int i = 0; i < n; i++) memset(&out[i], 0, sizeof out[i]);
    if (n <= 0) { if (t) *t = tm; return 0; }
    const SearchParams P = *(const SearchParams *)&enc->cfg.P;
    const int bd = in[0].bitDepth == 10 ? 10 : 8, bps = bd == 8 ? 1 : 2;
    (void)hipEventRecord(enc->ev[0], st);
    std::vector<Img> imgs(n);
    uint32_t shTotal[NSHAPES] = {0};
    std::vector<uint32_t> shStart((size_t)NSHAPES * (n + 1), 0), sbImg;
    uint32_t nSb = 0;
    bool any444 = false;
    int batchSs = -1;
    for (int i = 0; i < n; i++) {
        const av1mi_frame_in &f = in[i];
        if ((f.bitDepth == 10 ? 10 : 8) != bd) { snprintf(err, errlen, "frames of different bit depths in one batch"); return -1; }
        const int ss = (f.yuv444 && !f.mono) ? 0 : 1;   // chroma subsampling: 0 = 4:4:4, 1 = 4:2:0 (monochrome: no chroma)
        if (!f.mono) {
            if (batchSs < 0) batchSs = ss;
            else if (batchSs != ss) { snprintf(err, errlen, "frames of different chroma formats in one batch"); return -1; }
        }
        if (!ss) any444 = true;
        if (f.W <= 0 || f.H <= 0 || f.W > 65536 || f.H > 65536) { snprintf(err, errlen, "frame %d: bad size %dx%d", i, f.W, f.H); return -1; }
        Img &I = imgs[i];
        memset(&I, 0, sizeof I);
        I.W = f.W; I.H = f.H;
        I.MiCols = 2 * ((f.W + 7) >> 3); I.MiRows = 2 * ((f.H + 7) >> 3);
        I.sbCols = (I.MiCols + 15) >> 4; I.sbRows = (I.MiRows + 15) >> 4;
        if (f.aw < I.sbCols * 64 || f.ah < I.sbRows * 64) { snprintf(err, errlen, "frame %d: input planes not padded to 64", i); return -1; }
        I.aw = f.aw; I.ah = f.ah;
        I.ss = ss; I.mono = f.mono; I.bd = bd;
        I.qindex = f.qindex > 0 ? f.qindex : enc->cfg.qindex;
        if (I.qindex < 1 || I.qindex > 255) { snprintf(err, errlen, "frame %d: qindex %d outside 1..255", i, I.qindex); return -1; }
        I.qctx = I.qindex <= 20 ? 0 : I.qindex <= 60 ? 1 : I.qindex <= 120 ? 2 : 3;
        I.edgeFilter = enc->cfg.edgeFilter ? 1 : 0;
        frame_postfilter(enc->cfg, I);
        tile_layout(f.W, f.H, I.tileCols, I.tileRows, I.tileColStart, I.tileRowStart);
        for (int p = 0; p < 3; p++) { I.src[p] = f.dsrc[p]; I.spitch[p] = f.spitch[p]; }
        I.sbBase = nSb;
        nSb += (uint32_t)(I.sbCols * I.sbRows);
        for (int k = 0; k < I.sbCols * I.sbRows; k++) sbImg.push_back((uint32_t)i);
    }
    CoefPrefault pf;   // host outputs: their coefficient arrays are made while the device works
    if (!dev && !na) pf.start(imgs);
    uint64_t blkTotal = 0;
    for (int s = 0; s < NSHAPES; s++) {
        const uint64_t shapeBase = blkTotal;
        const bool en = s < BLOCK_INVALID && !(s >= 13 && s <= 15) && ((P.shapeMask >> s) & 1);
        for (int i = 0; i < n; i++) {
            Img &I = imgs[i];
            shStart[(size_t)s * (n + 1) + i] = shTotal[s];
            if (en) {
                const int w = 4 * bw4(s), h = 4 * bh4(s);
                I.shNbx[s] = (I.MiCols * 4 + w - 1) / w;
                I.shNby[s] = (I.MiRows * 4 + h - 1) / h;
            }
            I.shOff[s] = (uint32_t)(shapeBase + shTotal[s]);
            shTotal[s] += (uint32_t)(I.shNbx[s] * I.shNby[s]);
        }
        shStart[(size_t)s * (n + 1) + n] = shTotal[s];
        blkTotal += shTotal[s];
    }
    // ---- device arena ----
    Arena A;
    std::vector<uint8_t *> recMem(n), mapMem(n);
    size_t mapTotal = 0;
    for (int i = 0; i < n; i++) {
        const Img &I = imgs[i];
        A.add(recMem[i], ((size_t)I.aw * I.ah + 2 * (size_t)(I.aw >> I.ss) * (I.ah >> I.ss)) * bps);
        const size_t mi = (size_t)I.MiCols * I.MiRows, c1 = (size_t)(I.MiCols >> I.ss) * (I.MiRows >> I.ss);
        A.add(mapMem[i], 5 * mi + 2 * (mi + 2 * c1));
        mapTotal += 5 * mi + 2 * (mi + 2 * c1);
    }
    uint32_t *dStats, *dDone, *dQctl;
    unsigned long long *dFlops = nullptr;
    A.add(dStats, 8); A.add(dDone, nSb); A.add(dQctl, 2);
    if (enc->countWork) A.add(dFlops, 64);
    // the CDEF output planes (the reconstruction's geometry), zeroed like it
    std::vector<uint8_t *> cdefMem(n, nullptr);
    bool anyLf = false, anyCdef = false;
    for (int i = 0; i < n; i++) {
        const Img &I = imgs[i];
        anyLf = anyLf || I.lfLevel[0] || I.lfLevel[1];
        if (!I.cdefOn) continue;
        anyCdef = true;
        A.add(cdefMem[i], ((size_t)I.aw * I.ah + 2 * (size_t)(I.aw >> I.ss) * (I.ah >> I.ss)) * bps);
    }
    // everything so far starts each call at zero (reconstruction, maps, counters, done flags): one fill kernel; the
    // padding beyond the coded area is never reconstructed, so the returned planes are fully defined
    const size_t zeroBytes = A.total;
    Img *dImg; BlockRes *dBlk; uint32_t *dShStart, *dSbImg; uint16_t *dLeafCount; Leaf *dLeaves; float *dSbJ; SbOut *dSbOut;
    RecBlock *dOB; uint16_t *dOBT; RecTxb *dOT; int16_t *dOC;
    A.add(dImg, n); A.add(dBlk, blkTotal); A.add(dShStart, shStart.size()); A.add(dSbImg, nSb); A.add(dLeafCount, nSb);
    A.add(dLeaves, (size_t)nSb * MAX_LEAVES); A.add(dSbJ, nSb); A.add(dSbOut, nSb); A.add(dOB, (size_t)nSb * MAX_LEAVES);
    const uint32_t txbCap = any444 ? SB_TXB_CAP_444 : SB_TXB_CAP, coefCap = any444 ? SB_COEF_CAP_444 : SB_COEF_CAP;
    A.add(dOBT, (size_t)nSb * MAX_LEAVES); A.add(dOT, (size_t)nSb * txbCap); A.add(dOC, (size_t)nSb * coefCap);
    // wavefront step lists: per tile, step = (c - c0) + 2 (r - r0)
    std::vector<std::vector<uint32_t>> steps;
    for (int i = 0; i < n; i++) {
        const Img &I = imgs[i];
        for (int tr = 0; tr < I.tileRows; tr++)
            for (int tc = 0; tc < I.tileCols; tc++) {
                const int r0 = I.tileRowStart[tr] >> 4, r1 = (I.tileRowStart[tr + 1] + 15) >> 4;
                const int c0 = I.tileColStart[tc] >> 4, c1 = (I.tileColStart[tc + 1] + 15) >> 4;
                for (int r = r0; r < r1; r++)
                    for (int c = c0; c < c1; c++) {
                        const size_t s = (size_t)(c - c0) + 2 * (size_t)(r - r0);
                        if (steps.size() <= s) steps.resize(s + 1);
                        steps[s].push_back(I.sbBase + (uint32_t)(r * I.sbCols + c));
                    }
            }
    }
    std::vector<uint32_t> stepFlat, stepOff;
    for (auto &v : steps) { stepOff.push_back((uint32_t)stepFlat.size()); stepFlat.insert(stepFlat.end(), v.begin(), v.end()); }
    stepOff.push_back((uint32_t)stepFlat.size());
    if (stepFlat.size() != nSb) { snprintf(err, errlen, "internal: wavefront order covers %zu of %u superblocks", stepFlat.size(), nSb); return -1; }
    uint32_t *dSteps, *dLnMask = nullptr;
    PackOff *dPO;
    A.add(dSteps, stepFlat.size()); A.add(dPO, nSb);
    // deblocking work units per image (pass 0: each plane's rows x 4-sample columns; pass 1: 4-sample rows x columns)
    uint64_t *dDbStart = nullptr;
    std::vector<uint64_t> dbStart(2 * ((size_t)n + 1), 0);
    if (anyLf) {
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i < n; i++) {
                const Img &I = imgs[i];
                uint64_t u = 0;
                for (int p = 0; p < (I.mono ? 1 : 3); p++) {
                    const uint64_t rows = (uint64_t)(I.MiRows * 4) >> (p ? I.ss : 0), cols = (uint64_t)(I.MiCols * 4) >> (p ? I.ss : 0);
                    u += pass == 0 ? rows * (cols >> 2) : (rows >> 2) * cols;
                }
                dbStart[pass * ((size_t)n + 1) + i + 1] = dbStart[pass * ((size_t)n + 1) + i] + u;
            }
        A.add(dDbStart, dbStart.size());
    }
    uint8_t *dLnAct = nullptr;
    uint32_t lnN = std::min(nSb, enc->lnChunk);
    const size_t fixedBytes = A.total + (enc->ln ? (((size_t)nSb * 12 * 4 + 255) & ~(size_t)255) : 0) +
                              (enc->dpExport ? (((size_t)nSb * 255 * 4 + 255) & ~(size_t)255) : 0);
    if (na && enc->ln)   // the network's activations in what the scratch array has left: fewer superblocks per pass
        while (lnN > 1 && fixedBytes + ((ln_scratch_bytes(lnN) + 255) & ~(size_t)255) > na->bytes) lnN = (lnN + 1) / 2;
    if (enc->ln) { A.add(dLnMask, (size_t)nSb * 12); A.add(dLnAct, ln_scratch_bytes(lnN)); }
    float *dDpRel = nullptr;
    if (enc->dpExport) A.add(dDpRel, (size_t)nSb * 255);
    // ---- schedule knobs: the config's, or the autotuner's for this batch shape ----
    std::vector<uint64_t> sig = {(uint64_t)n, (uint64_t)bd, (uint64_t)(enc->ln != nullptr), enc->gen, enc->lnChunk};
    for (int i = 0; i < n; i++)
        sig.insert(sig.end(), {(uint64_t)imgs[i].W, (uint64_t)imgs[i].H, (uint64_t)imgs[i].aw, (uint64_t)imgs[i].ah, (uint64_t)imgs[i].ss,
                               (uint64_t)imgs[i].mono});
    int anS = std::max(1, enc->cfg.anStreams), sched = enc->cfg.schedule, cand = -1, occ = enc->cfg.queueOcc;
    av1mi_encoder::Tune *tu = nullptr;
    if (enc->cfg.tune) {
        tu = &enc->tunes[sig];
        const int k = tu->next < av1mi_encoder::NTUNE ? (cand = tu->next) : tu->best;
        anS = (k & 1) ? 4 : 1;
        sched = (k >> 1) == 2 ? 1 : 0;
        occ = (k >> 1) == 2 ? 0 : (k >> 1) + 1;
    }
    // closed-loop instance: two workgroups per CU when the batch keeps more superblocks ready per wavefront step than the
    // device has CUs (throughput), else one with the larger register budget (the latency of a superblock chain)
    const int wpe = occ ? occ : (uint64_t)nSb >= (uint64_t)std::max(1, enc->cus) * std::max<size_t>(1, steps.size()) ? 2 : 1;
    const int slots = enc->slots[bd == 10][wpe - 1] > 0 ? enc->slots[bd == 10][wpe - 1] : 1;
    const uint32_t qBlocks = enc->cfg.queueBlocks > 0 ? (uint32_t)enc->cfg.queueBlocks : (uint32_t)std::max(1, enc->cus * slots);
    // ---- device memory: stream-ordered per call, or the captured plan's (graph mode) ----
    bool capture = false;
    const bool graph = enc->cfg.graph && !na;
    if (na) {
        if (A.total > na->bytes) {
            snprintf(err, errlen, "av1mi_encode_noalloc: the batch needs %zu bytes of device scratch, the scratch array holds %zu",
                     A.total, na->bytes);
            return -1;
        }
        A.bind(na->mem);
    } else if (graph) {
        std::vector<uint64_t> key = sig;
        key.insert(key.end(), {(uint64_t)anS, (uint64_t)sched, (uint64_t)qBlocks, (uint64_t)A.total, (uint64_t)wpe});
        if (enc->plan.key != key) {
            drop_plan(enc);
            HCK(hipMalloc(&enc->plan.mem, A.total));
            enc->plan.key = key;
            capture = true;
        }
        A.bind(enc->plan.mem);
    } else {
        HCK(A.alloc(st));
    }
    tm.deviceBytes = A.total;
    for (int i = 0; i < n; i++) {
        Img &I = imgs[i];
        const size_t ysz = (size_t)I.aw * I.ah * bps, csz = (size_t)(I.aw >> I.ss) * (I.ah >> I.ss) * bps;
        I.rec[0] = recMem[i]; I.rec[1] = recMem[i] + ysz; I.rec[2] = recMem[i] + ysz + csz;
        I.pitch[0] = I.aw; I.pitch[1] = I.pitch[2] = I.aw >> I.ss;
        const size_t mi = (size_t)I.MiCols * I.MiRows, c1 = (size_t)(I.MiCols >> I.ss) * (I.MiRows >> I.ss);
        uint8_t *m = mapMem[i];
        I.mY = m; I.mUV = m + mi; I.mSkip = m + 2 * mi; I.mMiSize = m + 3 * mi; I.mTx = m + 4 * mi;
        uint8_t *c = m + 5 * mi;
        I.cLvl[0] = c; I.cLvl[1] = c + mi; I.cLvl[2] = c + mi + c1;
        I.cDc[0] = c + mi + 2 * c1; I.cDc[1] = c + 2 * mi + 2 * c1; I.cDc[2] = c + 2 * mi + 3 * c1;
        for (int p = 0; p < 3; p++) I.cdefOut[p] = nullptr;
        if (cdefMem[i]) { I.cdefOut[0] = cdefMem[i]; I.cdefOut[1] = cdefMem[i] + ysz; I.cdefOut[2] = cdefMem[i] + ysz + csz; }
    }
    Batch B;
    memset(&B, 0, sizeof B);
    B.nimg = n; B.img = dImg; B.T = enc->dT; B.blk = dBlk; B.shStart = dShStart; B.nSb = nSb; B.sbImg = dSbImg;
    B.leafCount = dLeafCount; B.leaves = dLeaves; B.sbJ = dSbJ; B.sbOut = dSbOut; B.oBlocks = dOB; B.oBlockTxb = dOBT;
    B.oTxb = dOT; B.oCoeffs = dOC; B.stats = dStats; B.txbCap = txbCap; B.coefCap = coefCap; B.dpRel = dDpRel;
    if (enc->ln) B.lnMask = dLnMask;
    B.flops = dFlops;
    // analysis hooks of this call: each shape's slice of the external offsets, and each shape's first record
    AnHook hooks[NSHAPES];
    uint32_t shBase[NSHAPES];
    for (int s = 0; s < NSHAPES; s++) {
        hooks[s] = AnHook{};
        if (enc->ext[s]) { hooks[s].ext = enc->ext[s]; hooks[s].extOff = enc->extOff + (size_t)s * n; hooks[s].extTopK = enc->extTopK; }
        if (enc->list[s]) { hooks[s].list = enc->list[s]; hooks[s].count = enc->listCount[s]; }
        shBase[s] = imgs[0].shOff[s];
    }
    const AnHooks anHooks{hooks, shBase};
    // the device part: zeroing, the pruning network (its time counts as analysis), the analysis shapes over anS
    // streams (fork / join by events), the DP, the closed loop.  Eager on the caller's stream, or captured once per batch
    // shape on the capture stream and replayed; no memset nodes, only kernels.
    auto device_part = [&](hipStream_t s, bool events) -> int {
        launch_zero(A.base, zeroBytes, s);
        if (events) (void)hipEventRecord(enc->ev[1], s);
        if (enc->ln && ln_masks(*enc->ln, dLnAct, dImg, dSbImg, nSb, enc->lnConf, lnN, dLnMask, s)) return -1;
        if (anS > 1) {
            hipStream_t sts[4] = {s, enc->aux[0], enc->aux[1], enc->aux[2]};
            if (hipEventRecord(enc->evFork, s) != hipSuccess) return -1;
            for (int k = 1; k < anS; k++) if (hipStreamWaitEvent(sts[k], enc->evFork, 0) != hipSuccess) return -1;
            launch_analysis_on(B, P, shTotal, bd, anHooks, sts, anS);
            for (int k = 1; k < anS; k++) {
                if (hipEventRecord(enc->evJoin[k - 1], sts[k]) != hipSuccess || hipStreamWaitEvent(s, enc->evJoin[k - 1], 0) != hipSuccess)
                    return -1;
            }
        } else {
            launch_analysis(B, P, shTotal, bd, anHooks, s);
        }
        if (events) (void)hipEventRecord(enc->ev[2], s);
        launch_dp(B, P, s);
        if (events) (void)hipEventRecord(enc->ev[3], s);
        if (sched == 1) {
            for (size_t k = 0; k + 1 < stepOff.size(); k++) launch_recon(B, P, dSteps + stepOff[k], stepOff[k + 1] - stepOff[k], bd, s);
        } else {
            RcQueue Q;
            Q.order = dSteps; Q.n = nSb; Q.head = dQctl; Q.abort = dQctl + 1; Q.done = dDone;
            Q.spinTicks = enc->clockHz * 30;   // no legitimate wait comes near 30 s
            launch_recon_queue(B, P, Q, std::min<uint32_t>(qBlocks, nSb), bd, wpe, s);
        }
        return hipGetLastError() == hipSuccess ? 0 : -1;
    };
    if (graph) {
        if (capture) {   // shape data once, then the capture
            HCK(hipMemcpy(dShStart, shStart.data(), shStart.size() * 4, hipMemcpyHostToDevice));
            HCK(hipMemcpy(dSbImg, sbImg.data(), (size_t)nSb * 4, hipMemcpyHostToDevice));
            HCK(hipMemcpy(dSteps, stepFlat.data(), stepFlat.size() * 4, hipMemcpyHostToDevice));
            HCK(hipStreamBeginCapture(enc->cap, hipStreamCaptureModeThreadLocal));
            const int r = device_part(enc->cap, false);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(enc->cap, &g);
            if (r || ec != hipSuccess || !g) {
                if (g) (void)hipGraphDestroy(g);
                drop_plan(enc);
                snprintf(err, errlen, "graph capture of the encode failed");
                return -1;
            }
            const hipError_t ei = hipGraphInstantiate(&enc->plan.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) { drop_plan(enc); snprintf(err, errlen, "HIP: %s (graph instantiate)", hipGetErrorString(ei)); return -1; }
        }
        HCK(hipMemcpyAsync(dImg, imgs.data(), n * sizeof(Img), hipMemcpyHostToDevice, st));   // pointers, qindex of this call
        (void)hipEventRecord(enc->ev[1], st);
        HCK(hipGraphLaunch(enc->plan.exec, st));
        (void)hipEventRecord(enc->ev[2], st);   // the whole device part counts as analysis in graph mode
        (void)hipEventRecord(enc->ev[3], st);
    } else {
        HCK(hipMemcpyAsync(dImg, imgs.data(), n * sizeof(Img), hipMemcpyHostToDevice, st));
        HCK(hipMemcpyAsync(dShStart, shStart.data(), shStart.size() * 4, hipMemcpyHostToDevice, st));
        HCK(hipMemcpyAsync(dSbImg, sbImg.data(), (size_t)nSb * 4, hipMemcpyHostToDevice, st));
        HCK(hipMemcpyAsync(dSteps, stepFlat.data(), stepFlat.size() * 4, hipMemcpyHostToDevice, st));
        if (device_part(st, true)) { snprintf(err, errlen, "HIP: launch failed (%s)", hipGetErrorString(hipGetLastError())); A.release(st); return -1; }
    }
    tm.schedule = (uint32_t)sched;
    tm.queueBlocks = sched == 1 ? 0 : std::min<uint32_t>(qBlocks, nSb);
    tm.anStreams = (uint32_t)anS;
    tm.graph = (uint32_t)graph;
    tm.tuned = tu && cand < 0 ? 1u : 0u;
    tm.queueOcc = sched == 1 ? 0u : (uint32_t)wpe;
    (void)hipEventRecord(enc->ev[4], st);
    tm.steps = (uint32_t)steps.size();
    tm.sbs = nSb;
    // post-filters on the finished reconstruction: deblocking (vertical, then horizontal edges), then CDEF into its own
    // planes (GpTimes::postfilter)
    (void)hipEventRecord(enc->ev[8], st);
    if (anyLf) {
        HCK(hipMemcpyAsync(dDbStart, dbStart.data(), dbStart.size() * 8, hipMemcpyHostToDevice, st));
        for (int pass = 0; pass < 2; pass++)
            launch_deblock(B, pass, dDbStart + pass * ((size_t)n + 1), dbStart[pass * ((size_t)n + 1) + n], bd, st);
    }
    if (anyCdef) launch_cdef(B, bd, st);
    HCK(hipGetLastError());
    (void)hipEventRecord(enc->ev[9], st);
    // ---- pack: prefix sums in decode order (tiles raster, SBs raster inside each tile) ----
    std::vector<SbOut> so(nSb);
    uint32_t stats[8], qctl[2];
    HCK(hipMemcpyAsync(so.data(), dSbOut, (size_t)nSb * sizeof(SbOut), hipMemcpyDeviceToHost, st));
    HCK(hipMemcpyAsync(stats, dStats, sizeof stats, hipMemcpyDeviceToHost, st));
    HCK(hipMemcpyAsync(qctl, dQctl, sizeof qctl, hipMemcpyDeviceToHost, st));
    unsigned long long flops[64] = {0};
    if (dFlops) HCK(hipMemcpyAsync(flops, dFlops, sizeof flops, hipMemcpyDeviceToHost, st));
    if (enc->dpExport) {
        enc->dpRel.resize((size_t)nSb * 255);
        HCK(hipMemcpyAsync(enc->dpRel.data(), dDpRel, enc->dpRel.size() * 4, hipMemcpyDeviceToHost, st));
    }
    if (enc->searchExport) {
        enc->exImg = imgs;
        enc->exBlk.resize(blkTotal); enc->exSbJ.resize(nSb); enc->exLeafCount.resize(nSb); enc->exLeaves.resize((size_t)nSb * MAX_LEAVES);
        HCK(hipMemcpyAsync(enc->exBlk.data(), dBlk, blkTotal * sizeof(BlockRes), hipMemcpyDeviceToHost, st));
        HCK(hipMemcpyAsync(enc->exSbJ.data(), dSbJ, (size_t)nSb * 4, hipMemcpyDeviceToHost, st));
        HCK(hipMemcpyAsync(enc->exLeafCount.data(), dLeafCount, (size_t)nSb * 2, hipMemcpyDeviceToHost, st));
        HCK(hipMemcpyAsync(enc->exLeaves.data(), dLeaves, (size_t)nSb * MAX_LEAVES * sizeof(Leaf), hipMemcpyDeviceToHost, st));
    }
    HCK(hipStreamSynchronize(st));
    if (qctl[1]) {
        snprintf(err, errlen, "closed-loop queue: a superblock dependency wait timed out (schedule aborted)");
        A.release(st);
        return -1;
    }
    std::vector<PackOff> po(nSb);
    std::vector<uint64_t> imgBlk(n + 1), imgTxb(n + 1), imgCoef(n + 1);
    uint64_t gBlk = 0, gTxb = 0, gCoef = 0;
    for (int i = 0; i < n; i++) {
        const Img &I = imgs[i];
        imgBlk[i] = gBlk; imgTxb[i] = gTxb; imgCoef[i] = gCoef;
        uint32_t rTxb = 0, rCoef = 0;
        for (int tr = 0; tr < I.tileRows; tr++)
            for (int tc = 0; tc < I.tileCols; tc++)
                for (int r = I.tileRowStart[tr]; r < I.tileRowStart[tr + 1]; r += 16)
                    for (int c = I.tileColStart[tc]; c < I.tileColStart[tc + 1]; c += 16) {
                        const uint32_t sb = I.sbBase + (uint32_t)((r >> 4) * I.sbCols + (c >> 4));
                        PackOff &o = po[sb];
                        o.gBlk = (uint32_t)gBlk; o.gTxb = (uint32_t)gTxb; o.gCoef = gCoef; o.rTxb = rTxb; o.rCoef = rCoef;
                        gBlk += so[sb].nBlocks; gTxb += so[sb].nTxb; gCoef += so[sb].nCoeffs;
                        rTxb += so[sb].nTxb; rCoef += so[sb].nCoeffs;
                    }
    }
    imgBlk[n] = gBlk; imgTxb[n] = gTxb; imgCoef[n] = gCoef;
    Arena R;
    RecBlock *rB; uint32_t *rBT; RecTxb *rT; int16_t *rC;
    // host outputs: the coefficients go down in sparse form (launch_sparse: spSpan spans of 512, spRun runs of 32 spans)
    const bool sparse = !na && !dev && gCoef > 0;
    const uint64_t spSpan = (gCoef + 511) / 512, spRun = (spSpan + 31) / 32;
    uint64_t *sMask = nullptr, *sOff = nullptr;
    uint32_t *sCnt = nullptr;
    int16_t *sVal = nullptr;
    if (na) {   // the caller's output buffers
        const av1mi_noalloc_out &o = *na->o;
        if (gBlk > o.cap_blocks || gTxb > o.cap_txb || gCoef > o.cap_coeffs) {
            snprintf(err, errlen, "av1mi_encode_noalloc: the batch has %llu blocks, %llu tx blocks and %llu coefficients, the output "
                     "buffers hold %llu, %llu and %llu", (unsigned long long)gBlk, (unsigned long long)gTxb, (unsigned long long)gCoef,
                     (unsigned long long)o.cap_blocks, (unsigned long long)o.cap_txb, (unsigned long long)o.cap_coeffs);
            return -1;
        }
        rB = (RecBlock *)o.blocks; rBT = o.block_txb; rT = (RecTxb *)o.txb; rC = o.coeffs;
    } else {
        R.add(rB, gBlk); R.add(rBT, gBlk); R.add(rT, gTxb); R.add(rC, gCoef);
        if (sparse) { R.add(sMask, spSpan * 8); R.add(sCnt, spRun); R.add(sOff, spRun + 1); R.add(sVal, gCoef); }
        HCK(R.alloc(st));
    }
    HCK(hipMemcpyAsync(dPO, po.data(), (size_t)nSb * sizeof(PackOff), hipMemcpyHostToDevice, st));
    (void)hipEventRecord(enc->ev[5], st);
    launch_pack(B, dPO, rB, rBT, rT, rC, st);
    HCK(hipGetLastError());
    (void)hipEventRecord(enc->ev[6], st);
    if (sparse) {   // masks and values of the nonzero coefficients, and their count to the host
        if (!enc->spTot) HCK(hipHostMalloc((void **)&enc->spTot, 64, hipHostMallocDefault));
        if (!enc->spEv) HCK(hipEventCreateWithFlags(&enc->spEv, hipEventDisableTiming));
        launch_sparse(rC, gCoef, sMask, sCnt, sOff, sVal, st);
        HCK(hipGetLastError());
        HCK(hipMemcpyAsync(enc->spTot, sOff + spRun, 8, hipMemcpyDeviceToHost, st));
        HCK(hipEventRecord(enc->spEv, st));
    }
    if (na) {   // the records are in the caller's buffers; the reconstruction goes to its planes, if it gave any
        for (int i = 0; i < n; i++) {
            const Img &I = imgs[i];
            av1mi_frame_out &o = out[i];
            o.width = I.W; o.height = I.H; o.base_q_idx = I.qindex; o.bit_depth = bd; o.chroma_format = I.ss ? 0 : 1;
            o.monochrome = I.mono;
            o.n_blocks = (uint32_t)(imgBlk[i + 1] - imgBlk[i]);
            o.n_txb = (uint32_t)(imgTxb[i + 1] - imgTxb[i]);
            o.n_coeffs = imgCoef[i + 1] - imgCoef[i];
            o.aw = I.aw; o.ah = I.ah; o.tile_cols = I.tileCols; o.tile_rows = I.tileRows;
            frame_params_fill(o, I);
            for (int p = 0; p < 3 && na->o->recon; p++) {
                void *dst = na->o->recon[3 * (size_t)i + p];
                if (!dst) continue;
                const int pw = p ? I.aw >> I.ss : I.aw, ph = p ? I.ah >> I.ss : I.ah;
                HCK(hipMemcpyAsync(dst, I.cdefOn ? I.cdefOut[p] : I.rec[p], (size_t)pw * ph * bps, hipMemcpyDeviceToDevice, st));
            }
        }
        (void)hipEventRecord(enc->ev[7], st);
        HCK(hipStreamSynchronize(st));
    } else {
        // records (and reconstructions) go to the host through the encoder's pinned staging buffer in one DMA per region,
        // then into each frame's own buffers on several host threads
        size_t reconBytes = 0;
        if (wantRecon)
            for (int i = 0; i < n; i++) reconBytes += ((size_t)imgs[i].aw * imgs[i].ah + 2 * (size_t)(imgs[i].aw >> imgs[i].ss) * (imgs[i].ah >> imgs[i].ss)) * bps;
        // the sparse form's value count: the copies below queue behind its kernels on this stream anyway, so waiting
        // for it here costs the device nothing and sizes the staging to what is downloaded
        uint64_t nnz = 0;
        if (sparse) {
            HCK(hipEventSynchronize(enc->spEv));
            nnz = *enc->spTot;
            if (nnz > gCoef) { snprintf(err, errlen, "internal: sparse coefficient count %llu", (unsigned long long)nnz); return -1; }
        }
        // staging: the records before the coefficients at their arena offsets; then, for the sparse form, the masks,
        // the runs' offsets and the values; then the reconstructions
        const size_t prefix = sparse ? (size_t)((const char *)rC - (const char *)R.base) : R.total;
        const size_t maskBytes = spSpan * 8 * sizeof(uint64_t);
        const size_t stMask = prefix, stOff = stMask + maskBytes, stVal = (stOff + (spRun + 1) * 8 + 255) & ~(size_t)255;
        const size_t recEnd = dev ? 0 : sparse ? (stVal + nnz * 2 + 255) & ~(size_t)255 : R.total;
        const size_t need = recEnd + reconBytes;
        if (need > enc->stageBytes) {
            HCK(hipStreamSynchronize(st));
            if (enc->stage) enc->oldStages.push_back(enc->stage);
            enc->stage = nullptr;
            const size_t cap = std::max(need + need / 4, 2 * enc->stageBytes);
            enc->stageBytes = 0;
            HCK(hipHostMalloc(&enc->stage, cap, hipHostMallocDefault));
            enc->stageBytes = cap;
        }
        char *stg = (char *)enc->stage;
        std::vector<size_t> segEnd;
        std::vector<hipEvent_t> segEv;
        auto segment = [&](size_t end) -> int {   // the download so far ends at staging offset end
            if (segEv.size() == enc->dlEv.size()) {
                hipEvent_t e;
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return -1;
                enc->dlEv.push_back(e);
            }
            segEv.push_back(enc->dlEv[segEv.size()]);
            segEnd.push_back(end);
            return hipEventRecord(segEv.back(), st) == hipSuccess ? 0 : -1;
        };
        constexpr size_t SEG = (size_t)64 << 20;
        auto download = [&](size_t at, const void *src, size_t bytes) -> hipError_t {   // to staging offset at, in segments
            for (size_t o = 0; o < bytes; o += SEG) {
                const hipError_t e = hipMemcpyAsync(stg + at + o, (const char *)src + o, std::min(SEG, bytes - o), hipMemcpyDeviceToHost, st);
                if (e != hipSuccess) return e;
                if (segment(at + std::min(o + SEG, bytes))) return hipErrorUnknown;
            }
            return hipSuccess;
        };
        if (!dev) {
            HCK(download(0, R.base, prefix));
            if (sparse) {
                HCK(download(stMask, sMask, maskBytes));
                HCK(download(stOff, sOff, (spRun + 1) * 8));
                HCK(download(stVal, sVal, nnz * 2));
            }
        }
        auto hostOf = [&](const void *dptr) { return stg + ((const char *)dptr - (const char *)R.base); };
        std::vector<std::tuple<void *, const void *, size_t>> jobs;
        size_t reconOff = recEnd;
        std::vector<int16_t *> coefDst;
        for (int i = 0; i < n; i++) {
            const Img &I = imgs[i];
            av1mi_frame_out &o = out[i];
            o.width = I.W; o.height = I.H; o.base_q_idx = I.qindex; o.bit_depth = bd; o.chroma_format = I.ss ? 0 : 1; o.monochrome = I.mono;
            o.n_blocks = (uint32_t)(imgBlk[i + 1] - imgBlk[i]);
            o.n_txb = (uint32_t)(imgTxb[i + 1] - imgTxb[i]);
            o.n_coeffs = imgCoef[i + 1] - imgCoef[i];
            o.aw = I.aw; o.ah = I.ah; o.tile_cols = I.tileCols; o.tile_rows = I.tileRows;
            frame_params_fill(o, I);
            if (!dev) {
                o.blocks = (av1mi_block *)malloc(std::max<size_t>(o.n_blocks, 1) * 16);
                o.block_txb = (uint32_t *)malloc(((size_t)o.n_blocks + 1) * 4);
                o.txb = (av1mi_txb *)out_alloc(std::max<size_t>(o.n_txb, 1) * 16);
                o.coeffs = sparse ? pf.take((size_t)i, o.n_coeffs) : nullptr;
                if (!o.coeffs) o.coeffs = (int16_t *)out_alloc(std::max<uint64_t>(o.n_coeffs, 1) * 2, sparse);
                jobs.emplace_back(o.blocks, hostOf(rB + imgBlk[i]), (size_t)o.n_blocks * 16);
                jobs.emplace_back(o.block_txb, hostOf(rBT + imgBlk[i]), (size_t)o.n_blocks * 4);
                jobs.emplace_back(o.txb, hostOf(rT + imgTxb[i]), (size_t)o.n_txb * 16);
                if (sparse) coefDst.push_back(o.coeffs);
                else jobs.emplace_back(o.coeffs, hostOf(rC + imgCoef[i]), (size_t)o.n_coeffs * 2);
            }
            if (wantRecon) {
                for (int p = 0; p < 3; p++) {
                    const int pw = p ? I.aw >> I.ss : I.aw, ph = p ? I.ah >> I.ss : I.ah;
                    const size_t b = (size_t)pw * ph * bps;
                    o.recon[p] = out_alloc(b);
                    o.recon_pitch[p] = pw;
                    HCK(hipMemcpyAsync(stg + reconOff, I.cdefOn ? I.cdefOut[p] : I.rec[p], b, hipMemcpyDeviceToHost, st));
                    jobs.emplace_back(o.recon[p], stg + reconOff, b);
                    reconOff += b;
                    if (segment(reconOff)) { snprintf(err, errlen, "HIP: download event"); return -1; }
                }
            }
        }
        (void)hipEventRecord(enc->ev[7], st);
        const int cr = parallel_copy(jobs, stg, segEnd, segEv);
        HCK(hipStreamSynchronize(st));
        if (cr) { snprintf(err, errlen, "internal: record copy"); return -1; }
        if (sparse && (sparse_expand((const uint64_t *)(stg + stMask), (const uint64_t *)(stg + stOff), (const int16_t *)(stg + stVal), gCoef,
                                     imgCoef, coefDst) ||
                       ((const uint64_t *)(stg + stOff))[spRun] != nnz)) {
            snprintf(err, errlen, "internal: sparse coefficients");
            return -1;
        }
        if (dev) {   // the packed arena passes to the caller as it is: frames back to back, frame-relative offsets
            memset(dev, 0, sizeof *dev);
            dev->n_frames = (uint32_t)n;
            dev->blk0 = (uint32_t *)malloc(((size_t)n + 1) * 4);
            dev->txb0 = (uint32_t *)malloc(((size_t)n + 1) * 4);
            dev->coef0 = (uint64_t *)malloc(((size_t)n + 1) * 8);
            for (int i = 0; i <= n; i++) { dev->blk0[i] = (uint32_t)imgBlk[i]; dev->txb0[i] = (uint32_t)imgTxb[i]; dev->coef0[i] = imgCoef[i]; }
            dev->blocks = (av1mi_block *)rB; dev->block_txb_rel = rBT; dev->txb = (av1mi_txb *)rT; dev->coeffs = rC;
            dev->pool = R.base;
            R.base = nullptr;
            R.owned = false;
        } else {
            for (int i = 0; i < n; i++) out[i].block_txb[out[i].n_blocks] = out[i].n_txb;
            R.release(st);
        }
    }
    A.release(st);
    float ms[7];
    for (int k = 0; k < 7; k++) (void)hipEventElapsedTime(&ms[k], enc->ev[k], enc->ev[k + 1]);
    (void)hipEventElapsedTime(&tm.postfilter, enc->ev[8], enc->ev[9]);
    tm.setup = ms[0]; tm.analysis = ms[1]; tm.dp = ms[2]; tm.recon = ms[3]; tm.pack = ms[5];
    tm.download = ms[6] + std::max(0.f, ms[4] - tm.postfilter);
    (void)hipEventElapsedTime(&tm.total, enc->ev[0], enc->ev[7]);
    if (cand >= 0) {   // autotuner: the device part's time of this candidate; after the last, keep the fastest
        tu->ms[cand] = ms[1] + ms[2] + ms[3];
        if (++tu->next == av1mi_encoder::NTUNE)
            for (int k = 1; k < av1mi_encoder::NTUNE; k++) if (tu->ms[k] < tu->ms[tu->best]) tu->best = k;
    }
    tm.blocks = stats[0]; tm.skipBlocks = stats[1]; tm.cflBlocks = stats[2]; tm.rangeFixes = stats[3];
    for (int k = 0; k < 32; k++) { tm.flopsAnalysis += flops[k]; tm.flopsRecon += flops[32 + k]; }
    if (t) *t = tm;
    return 0;
}

extern "C" {

int av1mi_encode(av1mi_encoder *enc, const av1mi_frame_in *in, int n, av1mi_frame_out *out, av1mi_times *t, void *stream, int wantRecon,
                 char *err, int errlen) {
    return encode_impl(enc, in, n, out, t, stream, wantRecon, nullptr, nullptr, err, errlen);
}
int av1mi_encode_noalloc(av1mi_encoder *enc, const av1mi_frame_in *in, int n, const av1mi_noalloc_out *o, av1mi_frame_out *out,
                         av1mi_times *t, void *stream, char *err, int errlen) {
    if (!o || (n > 0 && (!o->blocks || !o->block_txb || !o->txb || !o->coeffs))) {
        snprintf(err, errlen, "av1mi_encode_noalloc: no output buffers");
        return -1;
    }
    if (!enc->scratch) { snprintf(err, errlen, "av1mi_encode_noalloc: the device's scratch array is not available"); return -1; }
    const NoAlloc na{enc->scratch, enc->scratchBytes, o};
    return encode_impl(enc, in, n, out, t, stream, 0, nullptr, &na, err, errlen);
}
int av1mi_encode_dev(av1mi_encoder *enc, const av1mi_frame_in *in, int n, av1mi_frame_out *out, av1mi_times *t, void *stream,
                     int wantRecon, av1mi_dev_records *dev, char *err, int errlen) {
    if (!dev) { snprintf(err, errlen, "av1mi_encode_dev: no device record holder"); return -1; }
    memset(dev, 0, sizeof *dev);
    return encode_impl(enc, in, n, out, t, stream, wantRecon, dev, nullptr, err, errlen);
}
void av1mi_dev_records_free(av1mi_dev_records *dev, void *stream) {
    if (!dev) return;
    if (dev->pool) (void)hipFreeAsync(dev->pool, (hipStream_t)stream);
    free(dev->blk0); free(dev->txb0); free(dev->coef0);
    memset(dev, 0, sizeof *dev);
}

int av1mi_frames_to_10bit(av1mi_frame_in *f, int n, void **buf, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    *buf = nullptr;
    auto planeBytes = [](const av1mi_frame_in &x, int p) -> size_t {
        const int s = (p && !x.yuv444) ? 1 : 0;
        return ((size_t)(x.aw >> s) * (x.ah >> s) * 2 + 255) & ~(size_t)255;
    };
    size_t total = 0;
    for (int i = 0; i < n; i++)
        if (f[i].bitDepth != 10)
            for (int p = 0; p < (f[i].mono ? 1 : 3); p++) total += planeBytes(f[i], p);
    if (!total) return 0;
    uint8_t *d = nullptr;
    if (hipMallocAsync((void **)&d, total, st) != hipSuccess) return -1;
    size_t at = 0;
    for (int i = 0; i < n; i++) {
        av1mi_frame_in &x = f[i];
        if (x.bitDepth == 10) continue;
        for (int p = 0; p < (x.mono ? 1 : 3); p++) {
            const int s = (p && !x.yuv444) ? 1 : 0, w = x.aw >> s, h = x.ah >> s;
            uint16_t *o = (uint16_t *)(d + at);
            launch_to10((const uint8_t *)x.dsrc[p], x.spitch[p], o, w, w, h, p > 0, st);
            x.dsrc[p] = o;
            x.spitch[p] = w;
            at += planeBytes(x, p);
        }
        x.bitDepth = 10;
    }
    *buf = d;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void av1mi_frame_params_of(const av1mi_frame_out *o, int edgeFilter, int colorMatrix, av1mi_frame_params *fp) {
    memset(fp, 0, sizeof *fp);
    fp->version = 5;
    fp->width = o->width; fp->height = o->height;
    fp->monochrome = (uint8_t)o->monochrome;
    fp->chroma_format = (uint8_t)(o->monochrome ? 0 : o->chroma_format);
    fp->bit_depth = (uint8_t)o->bit_depth;
    fp->base_q_idx = (uint8_t)o->base_q_idx;
    fp->tile_cols_log2 = fp->tile_rows_log2 = -1;
    fp->enable_intra_edge_filter = (uint8_t)(edgeFilter ? 1 : 0);
    if (colorMatrix >= 0) {
        fp->color_description = 1; fp->color_primaries = 1; fp->transfer_characteristics = 13;
        fp->matrix_coefficients = (uint8_t)colorMatrix;
    } else if (colorMatrix == -2) {
        fp->color_description = 1;
        fp->color_primaries = fp->transfer_characteristics = fp->matrix_coefficients = 2;
    }
}

void av1mi_frame_out_free(av1mi_frame_out *o) {
    if (!o) return;
    free(o->blocks); free(o->block_txb); free(o->txb); free(o->coeffs);
    for (int p = 0; p < 3; p++) free(o->recon[p]);
    memset(o, 0, sizeof *o);
}

int av1mi_dev_alloc(void **ptr, size_t bytes) { return hipMalloc(ptr, bytes) == hipSuccess ? 0 : -1; }
int av1mi_dev_alloc_async(void **ptr, size_t bytes, void *stream) {
    return hipMallocAsync(ptr, bytes, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}
int av1mi_dev_free_async(void *ptr, void *stream) { return hipFreeAsync(ptr, (hipStream_t)stream) == hipSuccess ? 0 : -1; }
int av1mi_copy_h2d_async(void *dst, const void *src, size_t bytes, void *stream) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}
int av1mi_stream_create(void **stream) {
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -1;
    *stream = s;
    return 0;
}
int av1mi_stream_destroy(void *stream) { return hipStreamDestroy((hipStream_t)stream) == hipSuccess ? 0 : -1; }
int av1mi_stream_sync(void *stream) { return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : -1; }

int av1mi_compute_units(void) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return v;
}

size_t av1mi_frame_bytes(int W, int H, int bitDepth, int yuv444) {
    if (W <= 0 || H <= 0) return 0;
    const size_t bps = bitDepth == 10 ? 2 : 1;
    const size_t MiCols = 2 * (size_t)((W + 7) >> 3), MiRows = 2 * (size_t)((H + 7) >> 3);
    const size_t sbCols = (MiCols + 15) >> 4, sbRows = (MiRows + 15) >> 4, nSb = sbCols * sbRows;
    const size_t aw = sbCols * 64, ah = sbRows * 64, mi = MiCols * MiRows, c1 = yuv444 ? mi : mi / 4;
    const size_t txbCap = yuv444 ? SB_TXB_CAP_444 : SB_TXB_CAP, coefCap = yuv444 ? SB_COEF_CAP_444 : SB_COEF_CAP;
    size_t blocks = 0;
    for (int s = 0; s < BLOCK_INVALID; s++) {
        if (s >= 13 && s <= 15) continue;
        const size_t w = 4 * bw4(s), h = 4 * bh4(s);
        blocks += ((MiCols * 4 + w - 1) / w) * ((MiRows * 4 + h - 1) / h);
    }
    const size_t planes = yuv444 ? aw * ah * bps * 3 : aw * ah * bps * 3 / 2;
    const size_t arena = planes + 5 * mi + 2 * (mi + 2 * c1) + blocks * sizeof(BlockRes) +
                         nSb * (4 + 2 + MAX_LEAVES * sizeof(Leaf) + 4 + sizeof(SbOut) + MAX_LEAVES * (sizeof(RecBlock) + 2) +
                                txbCap * sizeof(RecTxb) + coefCap * 2 + 4 + sizeof(PackOff) + 4);
    const size_t records = nSb * (MAX_LEAVES * (sizeof(RecBlock) + 4) + txbCap * sizeof(RecTxb) + coefCap * 2);
    const size_t network = nSb * 48 + ln_scratch_bytes((uint32_t)std::min<size_t>(nSb, 65536));   // masks, activations
    return planes + arena + records + network + 64 * 1024;
}
int av1mi_dev_free(void *ptr) { return hipFree(ptr) == hipSuccess ? 0 : -1; }
int av1mi_copy_d2h(void *dst, const void *src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
int av1mi_copy_h2d(void *dst, const void *src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1; }
int av1mi_device_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }

int av1mi_fwd_matrix(int kind, int ln, float *out) {
    static DevTables *T = nullptr;
    if (!T) { T = (DevTables *)malloc(sizeof(DevTables)); build_tables(*T); }
    if (kind < 0 || kind > 2 || ln < 2 || ln > 6 || (kind == K_ADST && ln > 4) || (kind == K_IDT && ln > 5)) return 0;
    const int nn = 1 << (2 * ln);
    memcpy(out, T->fwd + T->fwdOff[kind][ln], nn * 4);
    return nn;
}
void av1mi_misc_tables(float *txg, float *txScale, float *avgY, float *avgPart, float *avgSkip0, float *lambda, float *lamS) {
    DevTables *T = (DevTables *)malloc(sizeof(DevTables));
    build_tables(*T);
    memcpy(txg, T->txg, sizeof T->txg); memcpy(txScale, T->txScale, sizeof T->txScale); memcpy(avgY, T->avgYMode, sizeof T->avgYMode);
    memcpy(avgPart, T->avgPart, sizeof T->avgPart); *avgSkip0 = T->avgSkip0;
    memcpy(lambda, T->lambda, sizeof T->lambda); memcpy(lamS, T->lamS, sizeof T->lamS);
    free(T);
}
int av1mi_cost_table(int qctx, float *out) {
    DevTables *T = (DevTables *)malloc(sizeof(DevTables));
    build_tables(*T);
    memcpy(out, &T->cost[qctx & 3], sizeof(CostTab));
    free(T);
    return (int)(sizeof(CostTab) / 4);
}

}  // extern "C"
