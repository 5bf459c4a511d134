// Compiles this library's field mappings (include/av1mi_bind.hpp) and C++ types (include/av1mi_gp.hpp) against the
// reference's own headers (encoder/src/gp.h, av1core/include/av1g/decisions.h), as a binding inside the reference tree
// would: identical layouts of the record entries and frame parameters, the same GpParams fields in the same order, and
// every mapped field carried by name; each unimplemented knob is refused.  Prints "binding ok" and exits 0.
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include "gp.h"
#include "av1g/decisions.h"
#include "av1mi_bind.hpp"
#include "av1mi_gp.hpp"

#define SAME_OFF(A, B, f) static_assert(offsetof(A, f) == offsetof(B, f), #A "." #f)
#define SAME_SIZE(A, B) static_assert(sizeof(A) == sizeof(B), #A " vs " #B)

SAME_SIZE(av1g_block, av1mi_block);
SAME_OFF(av1g_block, av1mi_block, mi_row); SAME_OFF(av1g_block, av1mi_block, mi_col); SAME_OFF(av1g_block, av1mi_block, bsize);
SAME_OFF(av1g_block, av1mi_block, skip); SAME_OFF(av1g_block, av1mi_block, tx_depth); SAME_OFF(av1g_block, av1mi_block, y_mode);
SAME_OFF(av1g_block, av1mi_block, angle_delta_y); SAME_OFF(av1g_block, av1mi_block, uv_mode);
SAME_OFF(av1g_block, av1mi_block, angle_delta_uv); SAME_OFF(av1g_block, av1mi_block, cfl_alpha_u);
SAME_OFF(av1g_block, av1mi_block, cfl_alpha_v); SAME_OFF(av1g_block, av1mi_block, filter_intra);
SAME_OFF(av1g_block, av1mi_block, cdef_idx); SAME_OFF(av1g_block, av1mi_block, sb_qindex);
SAME_SIZE(av1g_txb, av1mi_txb);
SAME_OFF(av1g_txb, av1mi_txb, coeff_off); SAME_OFF(av1g_txb, av1mi_txb, x4); SAME_OFF(av1g_txb, av1mi_txb, y4);
SAME_OFF(av1g_txb, av1mi_txb, plane); SAME_OFF(av1g_txb, av1mi_txb, tx_size); SAME_OFF(av1g_txb, av1mi_txb, tx_type);
SAME_OFF(av1g_txb, av1mi_txb, flags); SAME_OFF(av1g_txb, av1mi_txb, eob); SAME_OFF(av1g_txb, av1mi_txb, reserved);
SAME_SIZE(av1g_frame_params, av1mi_frame_params);
#define FP(f) SAME_OFF(av1g_frame_params, av1mi_frame_params, f)
FP(version); FP(width); FP(height); FP(monochrome); FP(base_q_idx); FP(tx_mode_select); FP(reduced_tx_set); FP(tile_cols_log2);
FP(tile_rows_log2); FP(delta_q_y_dc); FP(delta_q_u_dc); FP(delta_q_u_ac); FP(delta_q_v_dc); FP(delta_q_v_ac);
FP(enable_filter_intra); FP(enable_intra_edge_filter); FP(delta_q_present); FP(delta_q_res); FP(using_qmatrix); FP(qm_y);
FP(qm_u); FP(qm_v); FP(loop_filter_level); FP(loop_filter_sharpness); FP(enable_cdef); FP(cdef_damping_minus_3); FP(cdef_bits);
FP(cdef_y_pri); FP(cdef_y_sec); FP(cdef_uv_pri); FP(cdef_uv_sec); FP(allow_screen_content_tools); FP(chroma_format);
FP(bit_depth); FP(color_description); FP(color_primaries); FP(transfer_characteristics); FP(matrix_coefficients); FP(reserved);
SAME_SIZE(::GpParams, av1mi::GpParams);
#define GP(f) SAME_OFF(::GpParams, av1mi::GpParams, f)
GP(lambdaMul); GP(satdLambda); GP(rnd); GP(zbin); GP(rndC); GP(zbinC); GP(K); GP(Krecheck); GP(txSearch); GP(txSearchCL);
GP(deltas); GP(uvDeltas); GP(cfl); GP(uvK); GP(shapeMask); GP(partMask8); GP(partMaskN); GP(maxPart); GP(uvOverheadBits);
GP(splitBias); GP(rdoq); GP(lambdaCMul); GP(qmDist); GP(qmRnd); GP(smallKernel); GP(pruneVar); GP(pruneMinArea); GP(screenSub);
GP(filterIntra); GP(kLarge); GP(deltasMaxArea); GP(profile); GP(uvInDp); GP(txSizeSearch); GP(txAllCL); GP(poisonSmem);
GP(pruneFlat); GP(uvShortlist); GP(txShortlist);

static int fails = 0;
#define CHECK(c) do { if (!(c)) { printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); fails++; } } while (0)

int main() {
    ::GpParams g;
    memset(&g, 0, sizeof g);
    g.lambdaMul = 1.25f; g.satdLambda = 0.75f; g.rnd = 0.3f; g.zbin = 0.6f; g.rndC = 0.35f; g.zbinC = 0.65f;
    g.K = 5; g.Krecheck = 3; g.txSearch = 1; g.txSearchCL = 0; g.txAllCL = 1; g.deltas = 1; g.uvDeltas = 0; g.cfl = 1; g.uvK = 4;
    g.shapeMask = 0x3F0FFF; g.partMask8 = 0x7; g.partMaskN = 0xF0; g.maxPart = 32; g.uvOverheadBits = 2.5f; g.splitBias = 1.1f;
    g.lambdaCMul = 0.9f; g.smallKernel = 1; g.pruneMinArea = 1024; g.screenSub = 1;
    av1mi_params p;
    std::string err;
    CHECK(av1mi_bind::params_from(g, &p, err) == 0);
    CHECK(p.lambdaMul == 1.25f && p.satdLambda == 0.75f && p.rnd == 0.3f && p.zbin == 0.6f && p.rndC == 0.35f && p.zbinC == 0.65f);
    CHECK(p.K == 5 && p.Krecheck == 3 && p.txSearch == 1 && p.txSearchCL == 0 && p.txAllCL == 1 && p.deltas == 1 && p.uvDeltas == 0);
    CHECK(p.cfl == 1 && p.uvK == 4 && p.shapeMask == 0x3F0FFF && p.partMask8 == 0x7 && p.partMaskN == 0xF0 && p.maxPart == 32);
    CHECK(p.uvOverheadBits == 2.5f && p.splitBias == 1.1f && p.lambdaCMul == 0.9f);
    ::GpParams back;
    memset(&back, 0, sizeof back);
    av1mi_bind::params_to(p, back);
    CHECK(back.K == 5 && back.partMaskN == 0xF0 && back.splitBias == 1.1f && back.uvK == 4 && back.txAllCL == 1);
    // every knob the core lacks is refused, naming it
    struct { const char *name; void (*set)(::GpParams &); } bad[] = {
        {"rdoq", [](::GpParams &x) { x.rdoq = 1; }}, {"txSizeSearch", [](::GpParams &x) { x.txSizeSearch = 1; }},
        {"filterIntra", [](::GpParams &x) { x.filterIntra = 1; }}, {"qmDist", [](::GpParams &x) { x.qmDist = 1; }},
        {"qmRnd", [](::GpParams &x) { x.qmRnd = 1; }}, {"pruneVar", [](::GpParams &x) { x.pruneVar = 2.f; }},
        {"pruneFlat", [](::GpParams &x) { x.pruneFlat = 0.5f; }}, {"screenSub", [](::GpParams &x) { x.screenSub = 2; }},
        {"kLarge", [](::GpParams &x) { x.kLarge = 2; }}, {"deltasMaxArea", [](::GpParams &x) { x.deltasMaxArea = 256; }},
        {"uvInDp", [](::GpParams &x) { x.uvInDp = 1; }}, {"uvShortlist", [](::GpParams &x) { x.uvShortlist = 2; }},
        {"txShortlist", [](::GpParams &x) { x.txShortlist = 4; }},
    };
    for (auto &b : bad) {
        ::GpParams x = g;
        b.set(x);
        std::string e;
        CHECK(av1mi_bind::params_from(x, &p, e) == -1 && e.find(b.name) != std::string::npos);
    }
    // the frame parameters: every field by name gives the identical bytes
    av1mi_frame_params s;
    unsigned char *sb = (unsigned char *)&s;
    for (size_t i = 0; i < sizeof s; i++) sb[i] = (unsigned char)(7 * i + 3);
    memset(s.reserved, 0, sizeof s.reserved);
    s.version = 5; s.width = 4000; s.height = 3000;
    av1g_frame_params d;
    av1mi_bind::frame_params_to(s, d);
    CHECK(memcmp(&d, &s, offsetof(av1g_frame_params, reserved)) == 0);
    // the reference's defaults (pipeline.cpp:41-76) are av1mi::GpParams' defaults
    av1mi::GpParams dflt;
    CHECK(dflt.K == 4 && dflt.Krecheck == 2 && dflt.uvK == 2 && dflt.txAllCL == 0 && dflt.smallKernel == 1 && dflt.pruneMinArea == 1024);
    if (fails) return 1;
    printf("binding ok\n");
    return 0;
}
