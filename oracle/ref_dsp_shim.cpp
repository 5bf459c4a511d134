// Test infrastructure: C entry points over the reference's own header-only DSP (av1core/include/av1g/dsp.h,
// blockinfo.h) and its writer's symbol stream (av1core/src/internal.h: encode_with_hook's per-tile symbols; cdfctx.h:
// the default CDFs a tile starts from), compiled from the reference tree into oracle/_ref/ by oracle/Makefile.  The
// tests pin the oracle's integer restatement (orc_dsp.h) and its rate model against these.  Nothing in the product
// loads this library.
#include <stddef.h>
#include <string.h>
#include <vector>
#include "av1g/dsp.h"
#include "av1g/blockinfo.h"
#include "av1g/cdfctx.h"
#include "internal.h"

extern "C" {

unsigned ref_dequant_inverse(const int16_t *q, int txSz, int txType, int dcQ, int acQ, int bd, int32_t *res) {
    int32_t dq[1024];
    unsigned err = 0;
    if (bd == 10) { av1g_dequant<10>(q, dq, txSz, dcQ, acQ, nullptr); av1g_inv_txfm2d<10>(dq, txSz, txType, res, &err); }
    else { av1g_dequant<8>(q, dq, txSz, dcQ, acQ, nullptr); av1g_inv_txfm2d<8>(dq, txSz, txType, res, &err); }
    return err;
}

// above / left: AV1G_EDGE_LEN arrays (index AV1G_EDGE_OFF is sample 0), modified by the edge filter
void ref_predict(int mode, int delta, int edgeFilter, int filterType, int haveL, int haveA, int txSz, int x, int y, int maxX,
                 int maxY, uint16_t *above, uint16_t *left, int bd, int *pred) {
    const int lw = AV1G_T(Tx_Width_Log2)[txSz], lh = AV1G_T(Tx_Height_Log2)[txSz], w = 1 << lw, h = 1 << lh;
    if (bd == 10) {
        uint16_t dst[64 * 64];
        av1g_predict_intra_ex<10>(mode, delta, -1, edgeFilter, filterType, haveL, haveA, lw, lh, x, y, maxX, maxY,
                                  above + AV1G_EDGE_OFF, left + AV1G_EDGE_OFF, dst, w);
        for (int i = 0; i < w * h; i++) pred[i] = dst[i];
    } else {
        uint8_t a8[AV1G_EDGE_LEN], l8[AV1G_EDGE_LEN], dst[64 * 64];
        for (int i = 0; i < AV1G_EDGE_LEN; i++) { a8[i] = (uint8_t)above[i]; l8[i] = (uint8_t)left[i]; }
        av1g_predict_intra_ex<8>(mode, delta, -1, edgeFilter, filterType, haveL, haveA, lw, lh, x, y, maxX, maxY,
                                 a8 + AV1G_EDGE_OFF, l8 + AV1G_EDGE_OFF, dst, w);
        for (int i = 0; i < w * h; i++) pred[i] = dst[i];
    }
}

void ref_build_edges(const uint16_t *plane, int stride, int x, int y, int txSz, int haveL, int haveA, int haveAR, int haveBL,
                     int maxX, int maxY, int bd, uint16_t *above, uint16_t *left) {
    const int lw = AV1G_T(Tx_Width_Log2)[txSz], lh = AV1G_T(Tx_Height_Log2)[txSz];
    av1g_build_edges<10>(plane, stride, x, y, lw, lh, haveL, haveA, haveAR, haveBL, maxX, maxY, above + AV1G_EDGE_OFF,
                         left + AV1G_EDGE_OFF);
    (void)bd;
}

// The symbols the reference's writer codes for a record (av1core/src/encoder.cpp:1580-1584, TeeSink 197-214), all
// tiles concatenated in tile order: av1g_sym_pack words (cdf id = uint16 offset in av1g_cdfs, symbol, alphabet, kind).
// Returns the number of symbols (up to cap stored), or -1 with the writer's message in err.
long ref_tile_symbols(av1g_decision_record *rec, uint32_t *out, size_t cap, char *err, size_t errlen) {
    std::vector<std::vector<uint32_t>> ts;
    av1g_encode_opts o;
    memset(&o, 0, sizeof o);
    o.flags = AV1G_ENC_SKIP_RECON;
    o.seq_level_idx = -1;
    av1g_output res;
    const int rc = av1g::encode_with_hook(rec, &o, &res, nullptr, &ts);
    if (rc != AV1G_OK) {
        snprintf(err, errlen, "%s", res.error);
        return -1;
    }
    av1g_output_free(&res);
    size_t n = 0;
    for (const auto &t : ts)
        for (uint32_t w : t) {
            if (n < cap) out[n] = w;
            n++;
        }
    return (long)n;
}

// The default CDFs a tile of a frame with base_q_idx starts from (av1g_cdfs_init, cdfctx.h:98-155), as uint16 values
// at their cdf ids; returns the number of uint16 entries of av1g_cdfs.
size_t ref_default_cdfs(int base_q_idx, uint16_t *out) {
    av1g_cdfs c;
    av1g_cdfs_init(&c, base_q_idx);
    if (out) memcpy(out, &c, sizeof c);
    return sizeof c / 2;
}

// cdf ids (uint16 offsets in av1g_cdfs) and sizes in uint16 of the fields the coefficient syntax codes with, in this
// order: intra_tx_set1, intra_tx_set2, txb_skip, eob_pt_16 .. eob_pt_1024, eob_extra, dc_sign, coeff_base_eob,
// coeff_base, coeff_br (15 pairs)
int ref_coeff_cdf_fields(uint32_t *out) {
#define F_(f) out[k++] = (uint32_t)(offsetof(av1g_cdfs, f) / 2); out[k++] = (uint32_t)(sizeof(((av1g_cdfs *)0)->f) / 2);
    int k = 0;
    F_(intra_tx_set1) F_(intra_tx_set2) F_(txb_skip) F_(eob_pt_16) F_(eob_pt_32) F_(eob_pt_64) F_(eob_pt_128) F_(eob_pt_256)
    F_(eob_pt_512) F_(eob_pt_1024) F_(eob_extra) F_(dc_sign) F_(coeff_base_eob) F_(coeff_base) F_(coeff_br)
#undef F_
    return k / 2;
}

}
