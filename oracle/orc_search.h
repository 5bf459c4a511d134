// Oracle (test infrastructure only): CPU restatement of the reference's intra RD search -- tables, transform-block
// evaluation, open-loop analysis, partition DP, closed-loop reconstruction, pack -- with every float rounding
// explicit (see orc_search.cpp).  Used by tests/ as the parity checker and by bench.py as the CPU baseline.
#pragma once
#include <stdint.h>
#include <atomic>
#include <vector>
#include "orc_dsp.h"

namespace orc {

static const float INF = 3.0e38f;   // dev_common.cuh:15 (GP_INF)
static const int MAXK = 6;          // gp.h:19

// symbol costs in bits, -log2 of the default-CDF probabilities of one coefficient-CDF q-context (k_misc.cu:504-517)
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

struct Tables {
    // forward matrices F = inverse of the exact integer inverse transform (k_misc.cu:524-570): kind x log2 size
    std::vector<float> fwd[3][7];
    float txg[19][16];
    float txScale[19];
    CostTab cost[4];
    float avgYMode[13];
    float avgPart[5][10];
    float avgSkip0;
    float lambda[2][256], lamS[2][256];   // [bd == 10][qindex], lambdaMul = satdLambda = 1
};
const Tables &tables();

struct Params {   // GpParams (gp.h:29-68) restricted to the knobs of the plain search
    float lambdaMul = 1.f, satdLambda = 1.f;
    float rnd = 0.375f, zbin = 0.656f, rndC = 0.375f, zbinC = 0.656f;
    int K = 6, Krecheck = 6, txSearch = 1, txSearchCL = 1, txAllCL = 1, deltas = 1, uvDeltas = 1, cfl = 1, uvK = 6;
    int shapeMask = 0x3F1FFF, partMask8 = 0xF, partMaskN = 0x3FF, maxPart = 64;
    float uvOverheadBits = 1.5f, splitBias = 1.f, lambdaCMul = 1.f;
};

struct EvalIn {
    int txSz, txType, plane, dcQ, acQ;
    float rnd, zbin;
    int skipCtx, dcCtx, intraDir, bd;
    const CostTab *C;
    float resEnergy;
};
struct EvalOut { float D, R; int eob; };
// forward transform + dead-zone quantizer + distortion + exact coefficient rate of one tx block (dev_tx.cuh:271-316)
EvalOut eval_txb(const int16_t *res, const EvalIn &e, int16_t *q, float *coef = nullptr);
// the rate part alone: coefficient bits of the levels q (row-major tw x th) and their eob (dev_tx.cuh:47-165, 301-316)
float rate_of_levels(const int16_t *q, const EvalIn &e, int *eob);
float res_energy(const int16_t *res, int n);

struct BlockRes { float J; uint8_t cand[MAXK]; uint8_t ncand; uint8_t tx; };
struct Leaf { uint8_t bsize, x4, y4, pad; };

// av1g_block / av1g_txb of decisions.h:225-252 (same 16-byte layout)
struct RecBlock { uint16_t mi_row, mi_col; uint8_t bsize, skip, tx_depth, y_mode; int8_t angle_delta_y; uint8_t uv_mode;
                  int8_t angle_delta_uv, cfl_alpha_u, cfl_alpha_v; uint8_t filter_intra; int8_t cdef_idx; uint8_t sb_qindex; };
struct RecTxb { uint32_t coeff_off; uint16_t x4, y4; uint8_t plane, tx_size, tx_type, flags; uint16_t eob, reserved; };

// an external screener's candidate (GpExtCand, encoder/src/gp.h:136-143)
struct ExtCand { float J; uint8_t mode; int8_t delta; uint8_t tx, pad; };

struct Frame {
    int W, H, MiCols, MiRows, sbCols, sbRows, aw, ah, ssx, ssy, mono, bd, qindex, qctx, edgeFilter;
    int tileCols, tileRows, tileColStart[65], tileRowStart[65];
    std::vector<uint16_t> src[3], rec[3];
    int stride[3];
    std::vector<uint8_t> mY, mUV, mSkip, mMiSize, mTx, cLvl[3], cDc[3];
    int nbx[22], nby[22];
    std::vector<BlockRes> blk[22];
    std::vector<std::vector<Leaf>> leaves;      // per SB (raster), decode order
    std::vector<float> sbJ;
    // per-SB closed-loop outputs (SB-local txb offsets and coefficient offsets)
    std::vector<std::vector<RecBlock>> sbBlocks;
    std::vector<std::vector<uint32_t>> sbBlockTxb;
    std::vector<std::vector<RecTxb>> sbTxb;
    std::vector<std::vector<int16_t>> sbCoef;
    // packed record (decode order)
    std::vector<RecBlock> blocks;
    std::vector<uint32_t> blockTxb;
    std::vector<RecTxb> txb;
    std::vector<int16_t> coeffs;
    uint32_t stats[8];
    std::atomic<uint64_t> flops[2]{};   // work accounting: forward-transform FLOPs of the analysis / closed-loop evaluations
    // analysis hooks (GpBatch::extCand, gp.h:164-171; GpAnX::list, an_stage.h:23-27): per shape, the external candidates of each
    // block of the frame's grid (raster, extTopK each) replacing the screening, and which blocks are analysed at all
    const ExtCand *ext[22] = {};
    int extTopK = 0;
    const uint8_t *analyse[22] = {};
};

int tile_layout(int W, int H, int *cols, int *rows, int *colStart, int *rowStart);   // encoder.cpp:56-91 (min legal)
void frame_setup(Frame &F, int W, int H, int ssx, int mono, int bd, int qindex, int edgeFilter, const uint16_t *const src[3],
                 const int srcStride[3]);
void analysis(Frame &F, const Params &P, int shape);            // k_analysis.cuh:496-689
void partition_dp(Frame &F, const Params &P);                   // k_dp.cuh:72-297
void closed_loop_sb(Frame &F, const Params &P, int sb);         // k_recon.cuh:1692-1808
void pack(Frame &F);                                            // k_misc.cu:46-106, pipeline.cpp:878-904
void encode_frame(Frame &F, const Params &P, int threads);      // the whole path

}  // namespace orc
