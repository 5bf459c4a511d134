// The reference's C++ encoder API (encoder/src/pipeline.h:149-225: GpEncoder, gp_check_partitions,
// gp_frames_to_10bit, gp_dev_records_free) over this library's C ABI (include/av1mi.h), declared in
// include/av1mi_gp.hpp.  Host code only: every device step goes through the av1mi_* entry points.
#include <string.h>
#include <algorithm>
#include "../../include/av1mi_gp.hpp"
#include "../../include/av1mi_bind.hpp"

namespace av1mi {

namespace {

// GpConfig options outside this core: -1 and a message naming the first one set
int refuse_config(const GpConfig &c, std::string &err) {
    struct { bool bad; const char *what; } refused[] = {
        {c.qm != -1 || c.qmUV != -1, "qm / qmUV (quantizer matrices)"},
        {c.deltaq != 0 || c.deltaqStrength != 0, "deltaq (per-superblock qindex)"},
        {c.dqUDc || c.dqUAc || c.dqVDc || c.dqVAc, "dqUDc / dqUAc / dqVDc / dqVAc (chroma delta-q)"},
        {c.tuneIq || c.vb || c.fLf >= 0 || c.fCdef || c.ssimLambda, "tuneIq / vb / fLf / fCdef / ssimLambda (tune=iq tools)"},
        {c.tileMinMp > 0, "tileMinMp (requested tiling; this core codes minimum-legal tiles)"},
        {c.lossless != 0, "lossless"},
        {c.palette == 2, "palette 2 (palettes on every frame)"},
        {c.smallMp > 0, "smallMp (small-image preset)"},
        {c.an.topdown != 0 || c.an.prune41 >= 0, "an.topdown / an.prune41 (staged analysis)"},
    };
    for (const auto &r : refused)
        if (r.bad) {
            err = std::string("GpConfig.") + r.what + " is not implemented by this core";
            return -1;
        }
    return 0;
}

}  // namespace

av1mi_frame_out GpFrameOut::record() {
    av1mi_frame_out o;
    memset(&o, 0, sizeof o);
    o.width = fp.width; o.height = fp.height; o.base_q_idx = fp.base_q_idx; o.bit_depth = fp.bit_depth ? fp.bit_depth : 8;
    o.chroma_format = fp.chroma_format; o.monochrome = fp.monochrome;
    o.n_blocks = (uint32_t)blocks.size(); o.n_txb = (uint32_t)txb.size(); o.n_coeffs = coeffs.size();
    o.blocks = blocks.data(); o.block_txb = block_txb.data(); o.txb = txb.data(); o.coeffs = coeffs.data();
    o.tile_cols = tile_cols; o.tile_rows = tile_rows;
    o.fp = fp;
    return o;
}

GpEncoder::GpEncoder() {}
GpEncoder::~GpEncoder() {
    if (h_) av1mi_destroy(h_);
}

int GpEncoder::init(const GpConfig &cfg, std::string &err) {
    if (h_) { av1mi_destroy(h_); h_ = nullptr; }
    if (refuse_config(cfg, err)) return -1;
    av1mi_config c;
    memset(&c, 0, sizeof c);
    if (av1mi_bind::params_from(cfg.P, &c.P, err)) return -1;
    c.qindex = cfg.qindex;
    c.edgeFilter = cfg.edgeFilter;
    c.schedule = cfg.schedule; c.queueBlocks = cfg.queueBlocks; c.anStreams = cfg.anStreams; c.graph = cfg.graph; c.tune = cfg.tune;
    c.queueOcc = cfg.queueOcc;
    c.lfMode = cfg.lfMode; c.lfMul = cfg.lfMul; c.lfLevel = cfg.lfLevel;
    c.cdef = cfg.cdef; c.cdefDampingMinus3 = cfg.cdefDampingMinus3; c.cdefYPri = cfg.cdefYPri; c.cdefYSec = cfg.cdefYSec;
    c.cdefUVPri = cfg.cdefUVPri; c.cdefUVSec = cfg.cdefUVSec;
    char e[512] = {0};
    if (av1mi_create(&c, &h_, e, sizeof e)) { err = e; h_ = nullptr; return -1; }
    cfg_ = cfg;
    return 0;
}

int GpEncoder::encode(const std::vector<GpFrameIn> &in, std::vector<GpFrameOut> &out, GpTimes &t, std::string &err, hipStream_t st,
                      GpDevRecords *dev) {
    if (!h_) { err = "GpEncoder::encode before init"; return -1; }
    const int n = (int)in.size();
    std::vector<av1mi_frame_in> fi(n);
    for (int i = 0; i < n; i++) {
        const GpFrameIn &f = in[i];
        if (!f.sbQ.empty()) { err = "GpFrameIn.sbQ (per-superblock qindex) is not implemented by this core"; return -1; }
        av1mi_frame_in &x = fi[i];
        memset(&x, 0, sizeof x);
        x.W = f.W; x.H = f.H; x.mono = f.mono; x.yuv444 = f.yuv444; x.bitDepth = f.bitDepth;
        for (int p = 0; p < 3; p++) { x.dsrc[p] = f.dsrc[p]; x.spitch[p] = f.spitch[p]; }
        x.aw = f.aw; x.ah = f.ah; x.qindex = f.qindex;
    }
    std::vector<av1mi_frame_out> fo(std::max(n, 1));
    av1mi_times tm;
    av1mi_dev_records dr;
    char e[512] = {0};
    const int rc = dev ? av1mi_encode_dev(h_, fi.data(), n, fo.data(), &tm, st, cfg_.wantRecon ? 1 : 0, &dr, e, sizeof e)
                       : av1mi_encode(h_, fi.data(), n, fo.data(), &tm, st, cfg_.wantRecon ? 1 : 0, e, sizeof e);
    if (rc) { err = e; return -1; }
    out.clear();
    out.resize(n);
    for (int i = 0; i < n; i++) {
        av1mi_frame_out &o = fo[i];
        GpFrameOut &g = out[i];
        g.fp = o.fp;   // the encoder's frame decisions, with the frame's colour description
        av1mi_frame_params col;
        av1mi_frame_params_of(&o, cfg_.edgeFilter, in[i].colorMatrix, &col);
        g.fp.color_description = col.color_description; g.fp.color_primaries = col.color_primaries;
        g.fp.transfer_characteristics = col.transfer_characteristics; g.fp.matrix_coefficients = col.matrix_coefficients;
        g.tile_cols = o.tile_cols; g.tile_rows = o.tile_rows;
        if (!dev) {
            g.blocks.assign(o.blocks, o.blocks + o.n_blocks);
            g.block_txb.assign(o.block_txb, o.block_txb + o.n_blocks + 1);
            g.txb.assign(o.txb, o.txb + o.n_txb);
            g.coeffs.assign(o.coeffs, o.coeffs + o.n_coeffs);
        }
        if (cfg_.wantRecon && o.recon[0]) {   // cropped to the visible size
            const int bps = o.bit_depth > 8 ? 2 : 1, ss = (o.chroma_format == 1) ? 0 : 1;
            for (int p = 0; p < (o.monochrome ? 1 : 3); p++) {
                const int w = p ? (int)((o.width + ss) >> ss) : (int)o.width, h = p ? (int)((o.height + ss) >> ss) : (int)o.height;
                g.recon[p].resize((size_t)w * h * bps);
                for (int y = 0; y < h; y++)
                    memcpy(g.recon[p].data() + (size_t)y * w * bps, (const uint8_t *)o.recon[p] + (size_t)y * o.recon_pitch[p] * bps,
                           (size_t)w * bps);
            }
        }
        av1mi_frame_out_free(&o);
    }
    if (dev) {
        dev->fp.clear();
        for (int i = 0; i < n; i++) dev->fp.push_back(out[i].fp);
        dev->blk0.assign(dr.blk0, dr.blk0 + n + 1);
        dev->txb0.assign(dr.txb0, dr.txb0 + n + 1);
        dev->coef0.assign(dr.coef0, dr.coef0 + n + 1);
        dev->blocks = dr.blocks; dev->blockTxbRel = dr.block_txb_rel; dev->txb = dr.txb; dev->coeffs = dr.coeffs;
        dev->pool = dr.pool;
        free(dr.blk0); free(dr.txb0); free(dr.coef0);   // host copies taken; the device pool passes to dev
    }
    av1mi_bind::times_to(tm, t);
    return 0;
}

void gp_dev_records_free(GpDevRecords &R, hipStream_t st) {
    av1mi_dev_records d;
    memset(&d, 0, sizeof d);
    d.pool = R.pool;
    av1mi_dev_records_free(&d, st);
    R = GpDevRecords();
}

int gp_check_partitions(const GpParams &P, std::string &err) {
    av1mi_params p;
    if (av1mi_bind::params_from(P, &p, err)) return -1;
    char e[256] = {0};
    if (av1mi_check_partitions(&p, e, sizeof e)) { err = e; return -1; }
    return 0;
}

int gp_frames_to_10bit(std::vector<GpFrameIn> &frames, hipStream_t st, uint8_t **buf, std::string &err) {
    std::vector<av1mi_frame_in> fi(frames.size());
    for (size_t i = 0; i < frames.size(); i++) {
        av1mi_frame_in &x = fi[i];
        const GpFrameIn &f = frames[i];
        memset(&x, 0, sizeof x);
        x.W = f.W; x.H = f.H; x.mono = f.mono; x.yuv444 = f.yuv444; x.bitDepth = f.bitDepth;
        for (int p = 0; p < 3; p++) { x.dsrc[p] = f.dsrc[p]; x.spitch[p] = f.spitch[p]; }
        x.aw = f.aw; x.ah = f.ah;
    }
    void *b = nullptr;
    if (av1mi_frames_to_10bit(fi.data(), (int)fi.size(), &b, st)) { err = "gp_frames_to_10bit: HIP allocation or launch failed"; return -1; }
    for (size_t i = 0; i < frames.size(); i++) {
        frames[i].bitDepth = fi[i].bitDepth;
        for (int p = 0; p < 3; p++) { frames[i].dsrc[p] = (const uint8_t *)fi[i].dsrc[p]; frames[i].spitch[p] = fi[i].spitch[p]; }
    }
    *buf = (uint8_t *)b;
    return 0;
}

}  // namespace av1mi
