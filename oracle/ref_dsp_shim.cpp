// Test infrastructure: C entry points over the reference's own header-only DSP (av1core/include/av1g/dsp.h,
// blockinfo.h), compiled from the reference tree into oracle/_ref/ by oracle/Makefile.  The tests pin the oracle's
// integer restatement (orc_dsp.h) against these on random inputs.  Nothing in the product loads this library.
#include <stddef.h>
#include "av1g/dsp.h"
#include "av1g/blockinfo.h"

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

}
