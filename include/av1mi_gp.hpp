/* av1mi_gp.hpp - the reference's C++ encoder API over libav1mi.so (implemented in av1gpu_amd/csrc/gp_encoder.cpp).
 *
 * The types and calls of maddog107/av1gpu encoder/src/pipeline.h:149-225 and gp.h:29-68 for the intra RD path, with
 * the same names, fields, defaults and error convention (int 0 / -1 + std::string err), in namespace av1mi; the CUDA
 * stream becomes a hipStream_t.  A caller of the reference switches by including this header and linking libav1mi.so:
 *
 *   av1mi::GpEncoder enc;  std::string err;
 *   enc.init(cfg, err);                              // GpEncoder::init (pipeline.h:195)
 *   enc.encode(frames, outs, times, err, stream);    // GpEncoder::encode (pipeline.h:196-204)
 *
 * Options outside this core (quantizer matrices, delta-q, the post-filters, tune=iq, lossless, palettes on every frame,
 * per-image tiling, the small-image preset, staged analysis, the GpParams knobs named in av1mi_bind.hpp) are refused by
 * init with a message.  The front ends' options (crop / resize / trim / tone / gray, size limits, alpha, ycgcoRe) are
 * not part of the encoder and are not carried.
 */
#ifndef AV1MI_GP_HPP
#define AV1MI_GP_HPP

#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "av1mi.h"

namespace av1mi {

/* GpParams (gp.h:29-68): every field, the same order and meaning; defaults as GpConfig::GpConfig (pipeline.cpp:41-76) */
struct GpParams {
    float lambdaMul = 1.0f, satdLambda = 1.0f;
    float rnd = 0.375f, zbin = 0.656f, rndC = 0.375f, zbinC = 0.656f;
    int K = 4, Krecheck = 2, txSearch = 1, txSearchCL = 1, deltas = 1, uvDeltas = 1, cfl = 1, uvK = 2;
    int shapeMask = 0x3F1FFF, partMask8 = 0xF, partMaskN = 0x3FF, maxPart = 64;
    float uvOverheadBits = 1.5f, splitBias = 1.0f;
    int rdoq = 0;
    float lambdaCMul = 1.0f;
    int qmDist = 0, qmRnd = 0, smallKernel = 1;
    float pruneVar = 0.f;
    int pruneMinArea = 1024, screenSub = 1, filterIntra = 0, kLarge = 0, deltasMaxArea = 0, profile = 0, uvInDp = 0;
    int txSizeSearch = 0, txAllCL = 0, poisonSmem = 0;
    float pruneFlat = 0.f;
    int uvShortlist = 0, txShortlist = 0;
};

/* GpAnOpts (pipeline.h:36-41): staged analysis, refused unless off */
struct GpAnOpts {
    int topdown = 0;
    float prune41 = -1;
};

/* GpConfig (pipeline.h:44-105): the encoder's fields.  Besides the reference's, this core's launch knobs (none of
 * which changes a record): schedule / queueBlocks of the closed loop, anStreams / graph / tune (av1mi_config). */
struct GpConfig {
    int qindex = 100;
    GpParams P;
    GpAnOpts an;
    double smallMp = 0;
    GpParams Psmall;
    GpAnOpts anSmall;
    int qm = -1, qmUV = -1, deltaq = 0, deltaqRes = 2;
    float deltaqStrength = 0;
    int dqUDc = 0, dqUAc = 0, dqVDc = 0, dqVAc = 0;
    int edgeFilter = 0;
    int lfMode = 0;
    float lfMul = 1.0f;
    int lfLevel = 0;
    int cdef = 0, cdefDampingMinus3 = 0, cdefYPri = 0, cdefYSec = 0, cdefUVPri = 0, cdefUVSec = 0;
    bool verbose = false;
    bool wantRecon = true;
    int tuneIq = 0, vb = 0, vbOrder = 0, vbStrength = 100, fLf = -1, fCdef = 0, fDist = 0;
    float tileMinMp = 0;
    int tileColsLog2 = 1, tileRowsLog2 = 0;
    int ssimLambda = 0;
    int yuv444 = 0, depth = 8;
    int lossless = 0, losslessCfl = 1, llTileArea = 262144, ycgcoRe = 0, alpha = 1, alphaQindex = -1;
    int palette = 1;   /* 0 off, 1 frames that pass screen-content detection (photos never do: no palettes), 2 refused */
    int schedule = 0, queueBlocks = 0, anStreams = 1, graph = 0, tune = 0, queueOcc = 0;
};

/* GpFrameIn (pipeline.h:149-163) plus a per-frame qindex (0: the config's) */
struct GpFrameIn {
    int W = 0, H = 0, mono = 0;
    int yuv444 = 0;
    int bitDepth = 8;
    const uint8_t *dsrc[3] = {nullptr, nullptr, nullptr};
    int spitch[3] = {0, 0, 0};
    int aw = 0, ah = 0;
    std::vector<uint8_t> sbQ;   /* per-SB qindex targets (delta-q): refused when not empty */
    int jpegChroma = 0;         /* front ends only */
    int colorMatrix = -1;       /* CICP matrix of the planes (fp), -1 default, -2 unspecified */
    int qindex = 0;
};

/* GpFrameOut (pipeline.h:165-178): the DecisionRecord and the cropped reconstruction (Y W x H, U / V ((W + 1) / 2) x
 * ((H + 1) / 2), 4:4:4 W x H; 10-bit: 2 bytes per sample, native order) */
struct GpFrameOut {
    av1mi_frame_params fp;
    std::vector<av1mi_block> blocks;
    std::vector<uint32_t> block_txb;
    std::vector<av1mi_txb> txb;
    std::vector<int16_t> coeffs;
    std::vector<uint8_t> recon[3];
    int orientation = 1;
    bool dpNet = false;
    int tile_cols = 0, tile_rows = 0;
    /* the record as this library's writers take it (av1mi_write_obus / av1mi_write_avif); points into the vectors */
    av1mi_frame_out record();
};

/* GpTimes (pipeline.h:180-189) */
struct GpTimes {
    float setup = 0, analysis = 0, dp = 0, recon = 0, pack = 0, postfilter = 0, download = 0, total = 0;
    uint32_t steps = 0, sbs = 0, blocks = 0, skipBlocks = 0, cflBlocks = 0, rangeFixes = 0;
    uint32_t txSplit = 0;
    uint32_t dpRepaired = 0;
    size_t deviceBytes = 0;
    unsigned long long prof[16] = {0};
};

/* GpDevRecords (entropy_dev.h:13-34): device-resident records of a batch, frames back to back in decode order, all
 * inside pool (free with gp_dev_records_free) */
struct GpDevRecords {
    std::vector<av1mi_frame_params> fp;
    std::vector<uint32_t> blk0, txb0;
    std::vector<uint64_t> coef0;
    av1mi_block *blocks = nullptr;
    uint32_t *blockTxbRel = nullptr;
    av1mi_txb *txb = nullptr;
    int16_t *coeffs = nullptr;
    void *pool = nullptr;
};
void gp_dev_records_free(GpDevRecords &R, hipStream_t st);

class GpEncoder {
public:
    GpEncoder();
    ~GpEncoder();
    GpEncoder(const GpEncoder &) = delete;
    GpEncoder &operator=(const GpEncoder &) = delete;
    int init(const GpConfig &cfg, std::string &err);
    /* Encode a batch of frames (all on the GPU at once).  dev != nullptr: the records stay on the GPU (ownership
     * passes to dev); out[i] then only has fp (+ recon if wanted). */
    int encode(const std::vector<GpFrameIn> &in, std::vector<GpFrameOut> &out, GpTimes &t, std::string &err, hipStream_t st = 0,
               GpDevRecords *dev = nullptr);
    const GpConfig &config() const { return cfg_; }
    av1mi_encoder *handle() const { return h_; }

private:
    GpConfig cfg_;
    av1mi_encoder *h_ = nullptr;
};

/* gp_check_partitions (pipeline.h:216) */
int gp_check_partitions(const GpParams &P, std::string &err);
/* gp_frames_to_10bit (pipeline.h:218-225) */
int gp_frames_to_10bit(std::vector<GpFrameIn> &frames, hipStream_t st, uint8_t **buf, std::string &err);

}  // namespace av1mi

#endif
