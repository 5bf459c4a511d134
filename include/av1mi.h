/* av1mi.h - C ABI of the MI355X-native AV1 intra rate-distortion core (libav1mi.so).
 *
 * The drop-in boundary for the reference's GPU intra RD path.  The entry points mirror the reference's C++ encoder API
 * (maddog107/av1gpu encoder/src/pipeline.h:191-216: GpEncoder::init / encode / config, gp_check_partitions) and its
 * launch layer (encoder/src/kernels.h:11-82), flattened to plain C: an opaque encoder handle, POD structs, plain
 * pointers and sizes, a hipStream_t passed as void *.  Host C++ (include/av1mi_gp.hpp: the reference's GpEncoder /
 * GpConfig / GpParams / GpFrameIn / GpFrameOut / GpTimes / GpDevRecords, implemented in av1gpu_amd/csrc/gp_encoder.cpp)
 * and Python (av1gpu_amd/__init__.py) sit on top of these functions only.
 *
 * Output: for every frame, the DecisionRecord of the reference's av1core (av1core/include/av1g/decisions.h:225-322:
 * 16-byte av1g_block / av1g_txb entries in decode order, row-major int16 levels) and the exact reconstruction.
 * Errors: 0 = success, -1 = failure with a message in err (the reference's int 0 / -1 + std::string err convention).
 */
#ifndef AV1MI_H
#define AV1MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the DecisionRecord entries (decisions.h:225-252; identical 16-byte layouts) ---- */
typedef struct av1mi_block {
    uint16_t mi_row, mi_col;
    uint8_t bsize, skip, tx_depth, y_mode;
    int8_t angle_delta_y;
    uint8_t uv_mode;
    int8_t angle_delta_uv, cfl_alpha_u, cfl_alpha_v;
    uint8_t filter_intra;
    int8_t cdef_idx;
    uint8_t sb_qindex;
} av1mi_block;
typedef struct av1mi_txb {
    uint32_t coeff_off;
    uint16_t x4, y4;
    uint8_t plane, tx_size, tx_type, flags;
    uint16_t eob, reserved;
} av1mi_txb;

/* Frame-level decisions of a record: av1g_frame_params (decisions.h:269-313), the same 88-byte layout. */
typedef struct av1mi_frame_params {
    uint32_t version;
    uint32_t width, height;
    uint8_t monochrome, base_q_idx, tx_mode_select, reduced_tx_set;
    int8_t tile_cols_log2, tile_rows_log2;
    int8_t delta_q_y_dc, delta_q_u_dc, delta_q_u_ac, delta_q_v_dc, delta_q_v_ac;
    uint8_t enable_filter_intra, enable_intra_edge_filter;
    uint8_t delta_q_present, delta_q_res;
    uint8_t using_qmatrix, qm_y, qm_u, qm_v;
    uint8_t loop_filter_level[4];
    uint8_t loop_filter_sharpness;
    uint8_t enable_cdef, cdef_damping_minus_3, cdef_bits;
    uint8_t cdef_y_pri[8], cdef_y_sec[8], cdef_uv_pri[8], cdef_uv_sec[8];
    uint8_t allow_screen_content_tools;
    uint8_t chroma_format;
    uint8_t bit_depth;
    uint8_t color_description, color_primaries, transfer_characteristics, matrix_coefficients;
    uint8_t reserved[9];
} av1mi_frame_params;

/* Search knobs: the reference's GpParams (encoder/src/gp.h:29-68) fields this core implements, same names and meaning.
 * av1mi_default_params() gives the plain (maximum-effort) search: K = Krecheck = 6, every tx type in the closed loop,
 * angle deltas, chroma deltas, CfL, uvK = 6, all 19 shapes, all 10 partitions. */
typedef struct av1mi_params {
    float lambdaMul, satdLambda, rnd, zbin, rndC, zbinC;
    int K, Krecheck, txSearch, txSearchCL, txAllCL, deltas, uvDeltas, cfl, uvK, shapeMask, partMask8, partMaskN, maxPart;
    float uvOverheadBits, splitBias, lambdaCMul;
} av1mi_params;

/* GpConfig (pipeline.h:44-105) subset: qindex, the search knobs, enable_intra_edge_filter. */
typedef struct av1mi_config {
    int qindex;        /* base_q_idx 1..255 of frames that do not set their own */
    int edgeFilter;    /* enable_intra_edge_filter */
    av1mi_params P;
    /* closed-loop schedule (no counterpart in the reference, which launches once per wavefront step,
     * pipeline.cpp:815-827): 0 = persistent queue (default), 1 = one launch per wavefront step.  queueBlocks:
     * workgroups of the persistent kernel, 0 = every resident slot of the device.  Results do not depend on either. */
    int schedule;
    int queueBlocks;
    /* anStreams: the 19 analysis shape launches spread over 1..4 streams (0 = 1).  graph: capture the device part
     * of an encode (zeroing, pruning network, analysis, DP, closed loop) once per batch shape (image sizes, bit depth,
     * chroma format) as a hipGraph and replay it; per-image qindex and plane pointers may change between calls.
     * tune: the first 6 calls of each batch shape try anStreams 1 / 4 x closed loop (persistent queue at one or two
     * workgroups per CU, or one launch per step) and keep the fastest.  None of these changes a record. */
    int anStreams;
    int graph;
    int tune;
    /* post-filters (GpConfig, pipeline.h:60-66; frame levels as pipeline.cpp:960-986), decoder-exact on the GPU after
     * the closed loop; the returned reconstruction is the filtered frame.  lfMode: 0 off, 1 libaom's LPF_PICK_FROM_Q
     * level of each frame's qindex times lfMul, 2 the fixed lfLevel (0..63); cdef: 1 = one CDEF strength set for every
     * frame (cdef_bits 0): damping cdefDampingMinus3 + 3 (0..3), primary strengths 0..15, secondary 0, 1, 2 or 4. */
    int lfMode;
    float lfMul;
    int lfLevel;
    int cdef, cdefDampingMinus3, cdefYPri, cdefYSec, cdefUVPri, cdefUVSec;
    /* closed-loop workgroups per compute unit of the persistent schedule (no counterpart in the reference): 0 = by the
     * batch's parallelism (2 when it keeps more superblocks ready per wavefront step than the device has CUs), 1 or 2.
     * Results do not depend on it. */
    int queueOcc;
} av1mi_config;

/* GpFrameIn (pipeline.h:149-163): device planes padded (edge replicated) to aw x ah, multiples of 64; chroma
 * (aw >> 1) x (ah >> 1) for 4:2:0.  8-bit: uint8 samples; 10-bit: uint16 samples; pitches in samples.  qindex: this
 * frame's base_q_idx, or 0 for the config's (per-image qindex inside one batch is allowed). */
typedef struct av1mi_frame_in {
    int W, H, mono, yuv444, bitDepth;
    const void *dsrc[3];
    int spitch[3];
    int aw, ah;
    int qindex;
} av1mi_frame_in;

/* GpFrameOut (pipeline.h:165-178): host arrays owned by the library until av1mi_frame_out_free().  recon: the padded
 * reconstruction planes (aw x ah, chroma >> 1; 2 bytes per sample for 10-bit), NULL when not requested. */
typedef struct av1mi_frame_out {
    uint32_t width, height, base_q_idx, bit_depth, chroma_format, monochrome;
    uint32_t n_blocks, n_txb;
    uint64_t n_coeffs;
    av1mi_block *blocks;
    uint32_t *block_txb; /* n_blocks + 1 */
    av1mi_txb *txb;
    int16_t *coeffs;
    void *recon[3];
    int recon_pitch[3];
    int aw, ah;
    int tile_cols, tile_rows;
    /* the frame's decisions (GpFrameOut::fp, pipeline.h:166; set as pipeline.cpp:938-1007 with the default colour
     * description): sizes, format, base_q_idx, tiles, enable_intra_edge_filter, loop filter levels, CDEF */
    av1mi_frame_params fp;
} av1mi_frame_out;

/* GpTimes (pipeline.h:180-189): milliseconds from HIP events, and counts */
typedef struct av1mi_times {
    float setup, analysis, dp, recon, pack, download, total;
    uint32_t steps, sbs, blocks, skipBlocks, cflBlocks, rangeFixes;
    size_t deviceBytes;
    uint32_t schedule, queueBlocks;   /* the closed-loop schedule used (0 queue, 1 steps) and its workgroups */
    uint32_t anStreams, graph, tuned; /* analysis streams used; 1 = replayed / captured graph; 1 = the tuner's choice */
    /* with av1mi_set_work_count on: the algorithmic forward-transform FLOPs of the analysis (all shapes) and of the closed
     * loop (2 per fmaf of the column and row passes, 1 per multiply of identity kinds and of the row scale; 64-point
     * sizes keep 32 outputs); 0 otherwise */
    uint64_t flopsAnalysis, flopsRecon;
    float postfilter;   /* ms of deblocking + CDEF (GpTimes::postfilter) */
    uint32_t queueOcc;  /* persistent closed loop: the instance used, 1 or 2 workgroups per CU (0: per-step schedule) */
} av1mi_times;

typedef struct av1mi_encoder av1mi_encoder;

void av1mi_default_params(av1mi_params *p);
/* gp_check_partitions (pipeline.h:216): 0, or -1 when the shape / partition options leave some superblock no
 * partition of finite cost. */
int av1mi_check_partitions(const av1mi_params *p, char *err, int errlen);
/* GpEncoder::init (pipeline.h:195): selects the current HIP device, builds and uploads the tables. */
int av1mi_create(const av1mi_config *cfg, av1mi_encoder **out, char *err, int errlen);
void av1mi_destroy(av1mi_encoder *enc);
/* GpEncoder::config (pipeline.h:209) */
void av1mi_get_config(const av1mi_encoder *enc, av1mi_config *cfg);
/* GpEncoder::encode (pipeline.h:196-204): a batch of n frames (all of one bit depth) in flight at once on stream
 * (hipStream_t; NULL = default).  The input planes must stay valid until the call returns (it synchronises).
 * want_recon: also download the reconstruction. */
int av1mi_encode(av1mi_encoder *enc, const av1mi_frame_in *in, int n, av1mi_frame_out *out, av1mi_times *t, void *stream,
                 int want_recon, char *err, int errlen);
void av1mi_frame_out_free(av1mi_frame_out *out);
/* The frame parameters of an encoded frame as the reference sets them (pipeline.cpp:938-1007): version 5, sizes, bit
 * depth, chroma format, base_q_idx, TX_MODE_LARGEST, minimum-legal tiles, enable_intra_edge_filter; colorMatrix >= 0:
 * BT.709 primaries, sRGB transfer and that matrix; -2: unspecified (2/2/2); -1: the default colour description. */
void av1mi_frame_params_of(const av1mi_frame_out *out, int edge_filter, int color_matrix, av1mi_frame_params *fp);

/* Device-resident records (GpDevRecords, encoder/src/entropy_dev.h:13-34; GpEncoder::encode with dev != nullptr,
 * pipeline.cpp:1148-1168): the batch's records stay in device memory, frames back to back in decode order, one
 * allocation (pool) the caller owns until av1mi_dev_records_free.  blk0 / txb0 / coef0: host arrays of n + 1 frame bases;
 * block_txb_rel: each block's first tx block relative to its frame (no trailing entry); txb.coeff_off frame-relative. */
typedef struct av1mi_dev_records {
    uint32_t n_frames;
    uint32_t *blk0, *txb0;
    uint64_t *coef0;
    av1mi_block *blocks;
    uint32_t *block_txb_rel;
    av1mi_txb *txb;
    int16_t *coeffs;
    void *pool;
} av1mi_dev_records;
/* av1mi_encode with the records left on the device (dev): out[i] then carries the frame's sizes, counts and (with
 * want_recon) the reconstruction, and its record pointers are NULL. */
int av1mi_encode_dev(av1mi_encoder *enc, const av1mi_frame_in *in, int n, av1mi_frame_out *out, av1mi_times *t, void *stream,
                     int want_recon, av1mi_dev_records *dev, char *err, int errlen);
/* gp_dev_records_free (entropy_dev.h:43): releases dev's device allocation on stream and its host arrays. */
void av1mi_dev_records_free(av1mi_dev_records *dev, void *stream);

/* ---- encode without device allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  av1mi_encode_noalloc is av1mi_encode_dev with every internal buffer
 * (analysis records, leaves, superblock slots, maps, the pruning network's activations, the unpacked reconstruction) in
 * one file-scope device array of AV1MI_NOALLOC_SCRATCH_MIB MiB per device, reserved with the library's code and held
 * until the process ends; it calls no hipMalloc / hipFree / hipHostMalloc.  The caller passes the outputs: the packed
 * records in device buffers (frames back to back in batch order; block_txb each block's first tx block relative to its
 * frame, txb.coeff_off relative to its frame, as av1mi_dev_records), and optionally the reconstruction per frame.
 * out[i] receives the frame's sizes and counts (its pointers stay NULL).  A batch needing more scratch than the array, or
 * more records than the capacities, fails with -1 and a message before anything is written to the outputs.
 * Every call on one device shares that device's array: calls that overlap on one device (several host threads, or
 * several streams) race and must be serialised by the caller.  The encoder's graph option does not apply here (the
 * device part runs eagerly on stream). */
#define AV1MI_NOALLOC_SCRATCH_MIB 160
typedef struct av1mi_noalloc_out {
    av1mi_block *blocks;     /* device, cap_blocks entries */
    uint32_t *block_txb;     /* device, cap_blocks entries */
    av1mi_txb *txb;          /* device, cap_txb entries */
    int16_t *coeffs;         /* device, cap_coeffs entries */
    uint64_t cap_blocks, cap_txb, cap_coeffs;
    /* NULL, or 3 device plane pointers per frame (NULL entries skipped): the reconstruction, aw x ah luma and
     * (aw >> ss) x (ah >> ss) chroma samples (uint16_t at 10 bits), rows packed (pitch = width) */
    void *const *recon;
} av1mi_noalloc_out;
int av1mi_encode_noalloc(av1mi_encoder *enc, const av1mi_frame_in *in, int n, const av1mi_noalloc_out *o, av1mi_frame_out *out,
                         av1mi_times *t, void *stream, char *err, int errlen);

/* gp_frames_to_10bit (pipeline.h:218-225): the 8-bit frames of f become 10-bit device planes in place (Y10 =
 * round(Y8 * 1023 / 255), C10 = round((C8 - 128) * 1023 / 255) + 512), all in one stream-ordered allocation returned in
 * *buf (NULL if no frame was converted; free it with av1mi_dev_free_async after the encode).  The source planes must
 * stay valid until the work on stream is done. */
int av1mi_frames_to_10bit(av1mi_frame_in *f, int n, void **buf, void *stream);

/* Learned candidate pruning (no counterpart in the reference, whose pruning stages are heuristics:
 * encoder/src/k_leafprep.cu, an_stage.h, gp.h:136-143).  A BF16 CNN per 64x64 superblock predicts, per square
 * partition node, whether the analysis needs its children, its N x N / 2:1 shapes and its 4-way (4:1) shapes, and a
 * top-8 luma-mode shortlist; below
 * the confidence the search stays exhaustive (per node, and for the shortlist per superblock).  params:
 * av1mi_ln_param_count() floats (NULL disables).  The network runs in the encode call's stream-ordered memory,
 * AV1MI_LN_CHUNK superblocks at a time (environment, read here; default 65536, ~100 KB of activations each). */
int av1mi_set_shortlist(av1mi_encoder *enc, const float *params, float confidence, char *err, int errlen);
size_t av1mi_ln_param_count(void);
/* External screener (GpBatch::extCand / extOff / extTopK, encoder/src/gp.h:136-143 and 164-171, read by the analysis at
 * encoder/src/k_analysis.cuh:586-608): for each later encode, the blocks of shape s take their candidates from cand[s]
 * instead of the SATD screening: device memory, topK (1..32) av1mi_ext_cand per block, best first, the entry of block
 * `local` (the shape's raster order within its frame) of frame img at (off[s * n + img] + local) * topK, where n is
 * the number of frames of the call and off a device array of 22 * n offsets.  Entries with a mode of 13 or more are
 * skipped, the delta is kept for directional modes of blocks from 8x8 (values outside -3..3 count as 0), the first K
 * usable ones go to the RD; a block with none is screened as without the hook (the reference leaves that case
 * undefined).  cand NULL or all entries NULL: no hook.  The arrays must stay valid while encodes use them. */
typedef struct av1mi_ext_cand {
    float J;
    uint8_t mode;
    int8_t delta;
    uint8_t tx, pad;
} av1mi_ext_cand;
int av1mi_set_ext_candidates(av1mi_encoder *enc, const av1mi_ext_cand *const *cand, const int64_t *off, int topK, char *err,
                             int errlen);
/* Block lists (GpAnX::list / count, encoder/src/an_stage.h:23-27): for each later encode, shape s analyses only the blocks
 * list[s] names (device array of shape-global indices: the shape's blocks of frame 0, then frame 1, ..., each frame in
 * its raster order; count[s]: device pointer to the number of entries); its other blocks are not analysed (J = INF,
 * no candidate).  Refused when the shapes without a list leave some superblock no partition (av1mi_check_partitions on
 * them) and together with the learned pruning.  list NULL or all entries NULL: every block is analysed. */
int av1mi_set_block_lists(av1mi_encoder *enc, const uint32_t *const *list, const uint32_t *const *count, char *err, int errlen);
/* Training data for the pruning network (no counterpart in the reference): with on != 0, each later av1mi_encode keeps
 * the partition DP's per-node relative cost increases, superblocks in batch order, 255 floats each: (best cost without
 * the node's children - best) / best for the 85 square nodes, then the same without the node's N x N / 2:1 shapes,
 * then without its 4-way shapes.
 * av1mi_dp_export copies up to n floats of the last call and returns how many there are. */
int av1mi_set_dp_export(av1mi_encoder *enc, int on);
/* Work accounting (no counterpart in the reference): with on != 0 each later encode counts the forward-transform FLOPs
 * of every evaluation the analysis and the closed loop perform into av1mi_times.flopsAnalysis / flopsRecon (a few device
 * atomics per workgroup and superblock; records unchanged). */
int av1mi_set_work_count(av1mi_encoder *enc, int on);
size_t av1mi_dp_export(const av1mi_encoder *enc, float *out, size_t n);
/* Search internals of the last encode, for parity checks (the reference keeps them in device memory only: GpBlockRes,
 * GpLeaf and the per-SB cost of gp.h:112-132).  With on != 0 each later av1mi_encode keeps them:
 * av1mi_blockres_export: frame img's analysis records of one shape (block size), raster over the shape's grid, 12-byte
 * {float J; uint8 cand[6], ncand, tx} entries; av1mi_sbj_export: the partition DP's cost of each superblock (raster);
 * av1mi_leaves_export: per superblock the leaf count and 256 4-byte {bsize, x4, y4, pad} leaves in decode order.  Each
 * copies up to n (superblocks for the leaves) and returns how many there are. */
int av1mi_set_search_export(av1mi_encoder *enc, int on);
size_t av1mi_blockres_export(const av1mi_encoder *enc, int img, int shape, void *out, size_t n);
size_t av1mi_sbj_export(const av1mi_encoder *enc, int img, float *out, size_t n);
size_t av1mi_leaves_export(const av1mi_encoder *enc, int img, uint16_t *counts, void *leaves, size_t nsb);
void av1mi_ln_init_params(float *params, uint64_t seed);
/* Training on the device with this library's own forward (MFMA) and backward kernels. */
typedef struct av1mi_learner av1mi_learner;
int av1mi_ln_create(const float *params, av1mi_learner **out);
void av1mi_ln_destroy(av1mi_learner *l);
int av1mi_ln_get_params(av1mi_learner *l, float *params);
int av1mi_ln_forward(av1mi_learner *l, const uint8_t *tiles, const uint8_t *q, uint32_t n, float *logits);
int av1mi_ln_activations(av1mi_learner *l, int layer, uint32_t n, uint16_t *out);
int av1mi_ln_set_data(av1mi_learner *l, const uint8_t *tiles, const uint8_t *q, const float *label, const float *lmask, uint32_t n);
int av1mi_ln_gradient(av1mi_learner *l, uint32_t first, uint32_t n, float *loss, float *grads);
int av1mi_ln_train_step(av1mi_learner *l, uint32_t first, uint32_t n, float lr, float weight_decay, float *loss);

/* The bitstream of a record (the reference's host writer, av1core/src/encoder.cpp:1072-1659 with avif.cpp):
 * av1mi_write_obus gives the sequence header and frame OBUs of the still picture (default initial CDFs adapted per
 * symbol; the loop filter and CDEF parameters of rec->fp when its version is set, each superblock's cdef_idx from its
 * first coded block),
 * av1mi_write_avif the same wrapped in an AVIF file.  edge_filter: enable_intra_edge_filter of the encode.  *data is
 * malloc'd (release with av1mi_buffer_free).  0, or -1 with a message when the record is not codable. */
int av1mi_write_obus(const av1mi_frame_out *rec, int edge_filter, void **data, size_t *size, char *err, int errlen);
int av1mi_write_avif(const av1mi_frame_out *rec, int edge_filter, void **data, size_t *size, char *err, int errlen);
void av1mi_buffer_free(void *p);

/* Device-buffer helpers for callers without their own allocator (hipMalloc / hipFree / hipMemcpy). */
int av1mi_dev_alloc(void **ptr, size_t bytes);
int av1mi_dev_free(void *ptr);
int av1mi_copy_h2d(void *dst, const void *src, size_t bytes);
int av1mi_copy_d2h(void *dst, const void *src, size_t bytes);
int av1mi_device_sync(void);
/* Stream-ordered variants and non-blocking streams, for callers that overlap several encodes on one device (the
 * serving loop): allocation / free from the device's stream-ordered pool, asynchronous host-to-device copy. */
int av1mi_dev_alloc_async(void **ptr, size_t bytes, void *stream);
int av1mi_dev_free_async(void *ptr, void *stream);
int av1mi_copy_h2d_async(void *dst, const void *src, size_t bytes, void *stream);
int av1mi_stream_create(void **stream);
int av1mi_stream_destroy(void *stream);
int av1mi_stream_sync(void *stream);
/* Upper bound of the device bytes one frame of W x H occupies during av1mi_encode (its padded input planes, the
 * encode's arena share including the pruning network's activations and masks, and its compacted records): the
 * admission unit of a serving loop. */
size_t av1mi_frame_bytes(int W, int H, int bitDepth, int yuv444);
/* compute units of the current device (0 if there is none) */
int av1mi_compute_units(void);

/* Seeded synthetic test image (the benchmark generator): W x H planes, chroma ((W + ssx) >> ssx) x ((H + ssx) >> ssx);
 * bd 8: uint8 samples, bd 10: uint16; strides in samples.  Host memory. */
int av1mi_synth_image(uint64_t seed, int W, int H, int ssx, int bd, void *y, void *u, void *v, const int *strides);

/* Host copies of the device tables, for the oracle comparison tests: forward matrix (kind 0 DCT / 1 ADST / 2 identity,
 * log2 size) -> n * n floats or 0; the other tables as in the oracle's orc_misc_tables / orc_cost_table. */
int av1mi_fwd_matrix(int kind, int ln, float *out);
void av1mi_misc_tables(float *txg, float *txScale, float *avgY, float *avgPart, float *avgSkip0, float *lambda, float *lamS);
int av1mi_cost_table(int qctx, float *out);

#ifdef __cplusplus
}
#endif
#endif
