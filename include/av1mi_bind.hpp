/* av1mi_bind.hpp - field-by-field mapping between the reference's host types and this library's C ABI (av1mi.h).
 *
 * Header-only templates over the type of the search knobs, the frame parameters and the times, so the same code serves
 * this library's own C++ API (av1mi_gp.hpp, whose types carry the reference's field names) and a binding inside the
 * reference tree (GpParams of encoder/src/gp.h:29-68, av1g_frame_params of av1core/include/av1g/decisions.h:269-313,
 * GpTimes of encoder/src/pipeline.h:180-189).  Every field is copied by name; knobs this core does not implement are
 * refused with a message naming them instead of being ignored.
 */
#ifndef AV1MI_BIND_HPP
#define AV1MI_BIND_HPP

#include <string.h>
#include <string>
#include "av1mi.h"

namespace av1mi_bind {

/* GpParams -> av1mi_params.  Refused (-1 and err): rdoq, txSizeSearch, filterIntra, qmDist, qmRnd, pruneVar > 0,
 * pruneFlat > 0, screenSub > 1, kLarge > 0, deltasMaxArea > 0, uvInDp, uvShortlist, txShortlist.  Accepted without
 * effect on the decisions: smallKernel (the reference's thread mapping of its analysis kernels for small shapes; this
 * core has one analysis whose results follow its float contract), pruneMinArea (read only with pruneVar), profile and
 * poisonSmem (instrumentation of the reference's kernels). */
template <class GpParamsT>
int params_from(const GpParamsT &g, av1mi_params *p, std::string &err) {
    struct { bool bad; const char *what; } refused[] = {
        {g.rdoq != 0, "rdoq (greedy RDOQ in the closed loop)"},
        {g.txSizeSearch != 0, "txSizeSearch (closed-loop tx-size search, TX_MODE_SELECT)"},
        {g.filterIntra != 0, "filterIntra (filter-intra candidates)"},
        {g.qmDist != 0, "qmDist (quantizer-matrix weighted distortion)"},
        {g.qmRnd != 0, "qmRnd (quantizer-matrix scaled rounding)"},
        {g.pruneVar > 0.f, "pruneVar (variance pruning of the analysis)"},
        {g.pruneFlat > 0.f, "pruneFlat (flat-region pruning of the analysis)"},
        {g.screenSub > 1, "screenSub 2 (subsampled SATD screening)"},
        {g.kLarge > 0, "kLarge (separate K for large shapes)"},
        {g.deltasMaxArea > 0, "deltasMaxArea (angle deltas for small shapes only)"},
        {g.uvInDp != 0, "uvInDp (chroma term in the analysis J)"},
        {g.uvShortlist != 0, "uvShortlist (open-loop chroma shortlists)"},
        {g.txShortlist != 0, "txShortlist (open-loop tx-type shortlists)"},
    };
    for (const auto &r : refused)
        if (r.bad) {
            err = std::string("GpParams.") + r.what + " is not implemented by this core; leave it at 0";
            return -1;
        }
    memset(p, 0, sizeof *p);
    p->lambdaMul = g.lambdaMul; p->satdLambda = g.satdLambda;
    p->rnd = g.rnd; p->zbin = g.zbin; p->rndC = g.rndC; p->zbinC = g.zbinC;
    p->K = g.K; p->Krecheck = g.Krecheck; p->txSearch = g.txSearch; p->txSearchCL = g.txSearchCL; p->txAllCL = g.txAllCL;
    p->deltas = g.deltas; p->uvDeltas = g.uvDeltas; p->cfl = g.cfl; p->uvK = g.uvK;
    p->shapeMask = g.shapeMask; p->partMask8 = g.partMask8; p->partMaskN = g.partMaskN; p->maxPart = g.maxPart;
    p->uvOverheadBits = g.uvOverheadBits; p->splitBias = g.splitBias; p->lambdaCMul = g.lambdaCMul;
    return 0;
}

/* av1mi_params -> GpParams (the fields av1mi_params carries; the others keep their values) */
template <class GpParamsT>
void params_to(const av1mi_params &p, GpParamsT &g) {
    g.lambdaMul = p.lambdaMul; g.satdLambda = p.satdLambda;
    g.rnd = p.rnd; g.zbin = p.zbin; g.rndC = p.rndC; g.zbinC = p.zbinC;
    g.K = p.K; g.Krecheck = p.Krecheck; g.txSearch = p.txSearch; g.txSearchCL = p.txSearchCL; g.txAllCL = p.txAllCL;
    g.deltas = p.deltas; g.uvDeltas = p.uvDeltas; g.cfl = p.cfl; g.uvK = p.uvK;
    g.shapeMask = p.shapeMask; g.partMask8 = p.partMask8; g.partMaskN = p.partMaskN; g.maxPart = p.maxPart;
    g.uvOverheadBits = p.uvOverheadBits; g.splitBias = p.splitBias; g.lambdaCMul = p.lambdaCMul;
}

/* av1mi_frame_params -> av1g_frame_params, field by field */
template <class FrameParamsT>
void frame_params_to(const av1mi_frame_params &s, FrameParamsT &d) {
    memset(&d, 0, sizeof d);
    d.version = s.version; d.width = s.width; d.height = s.height;
    d.monochrome = s.monochrome; d.base_q_idx = s.base_q_idx; d.tx_mode_select = s.tx_mode_select; d.reduced_tx_set = s.reduced_tx_set;
    d.tile_cols_log2 = s.tile_cols_log2; d.tile_rows_log2 = s.tile_rows_log2;
    d.delta_q_y_dc = s.delta_q_y_dc; d.delta_q_u_dc = s.delta_q_u_dc; d.delta_q_u_ac = s.delta_q_u_ac;
    d.delta_q_v_dc = s.delta_q_v_dc; d.delta_q_v_ac = s.delta_q_v_ac;
    d.enable_filter_intra = s.enable_filter_intra; d.enable_intra_edge_filter = s.enable_intra_edge_filter;
    d.delta_q_present = s.delta_q_present; d.delta_q_res = s.delta_q_res;
    d.using_qmatrix = s.using_qmatrix; d.qm_y = s.qm_y; d.qm_u = s.qm_u; d.qm_v = s.qm_v;
    for (int k = 0; k < 4; k++) d.loop_filter_level[k] = s.loop_filter_level[k];
    d.loop_filter_sharpness = s.loop_filter_sharpness;
    d.enable_cdef = s.enable_cdef; d.cdef_damping_minus_3 = s.cdef_damping_minus_3; d.cdef_bits = s.cdef_bits;
    for (int k = 0; k < 8; k++) {
        d.cdef_y_pri[k] = s.cdef_y_pri[k]; d.cdef_y_sec[k] = s.cdef_y_sec[k];
        d.cdef_uv_pri[k] = s.cdef_uv_pri[k]; d.cdef_uv_sec[k] = s.cdef_uv_sec[k];
    }
    d.allow_screen_content_tools = s.allow_screen_content_tools; d.chroma_format = s.chroma_format; d.bit_depth = s.bit_depth;
    d.color_description = s.color_description; d.color_primaries = s.color_primaries;
    d.transfer_characteristics = s.transfer_characteristics; d.matrix_coefficients = s.matrix_coefficients;
}

/* av1mi_times -> GpTimes (ms and counts; fields without a counterpart stay 0) */
template <class GpTimesT>
void times_to(const av1mi_times &s, GpTimesT &t) {
    t.setup = s.setup; t.analysis = s.analysis; t.dp = s.dp; t.recon = s.recon; t.pack = s.pack; t.download = s.download;
    t.total = s.total; t.postfilter = s.postfilter;
    t.steps = s.steps; t.sbs = s.sbs; t.blocks = s.blocks; t.skipBlocks = s.skipBlocks; t.cflBlocks = s.cflBlocks;
    t.rangeFixes = s.rangeFixes; t.txSplit = 0; t.dpRepaired = 0; t.deviceBytes = s.deviceBytes;
}

}  // namespace av1mi_bind

#endif
