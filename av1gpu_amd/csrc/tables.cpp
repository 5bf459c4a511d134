// Host construction of the device tables (DevTables): forward transform matrices, distortion gains, output scales,
// symbol costs, open-loop averages and the per-qindex lambdas.  The recipe follows the reference's gp_init_tables
// (encoder/src/k_misc.cu:477-644) and gp_lambda (dev_common.cuh:99-107); everything is computed in double and rounded
// to float once, so the values do not depend on the device.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "av1m_common.h"

namespace av1m {

namespace {

// Gauss-Jordan inversion with partial pivoting (first largest pivot), row-major n x n
std::vector<double> gj_inverse(std::vector<double> a, int n) {
    std::vector<double> r((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) r[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int k = c + 1; k < n; k++)
            if (fabs(a[(size_t)k * n + c]) > fabs(a[(size_t)piv * n + c])) piv = k;
        if (piv != c)
            for (int k = 0; k < n; k++) {
                std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
                std::swap(r[(size_t)c * n + k], r[(size_t)piv * n + k]);
            }
        const double d = a[(size_t)c * n + c];
        for (int k = 0; k < n; k++) {
            a[(size_t)c * n + k] /= d;
            r[(size_t)c * n + k] /= d;
        }
        for (int row = 0; row < n; row++) {
            if (row == c) continue;
            const double f = a[(size_t)row * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; k++) {
                a[(size_t)row * n + k] -= f * a[(size_t)c * n + k];
                r[(size_t)row * n + k] -= f * r[(size_t)c * n + k];
            }
        }
    }
    return r;
}

// bits of each symbol of nc consecutive CDFs of nsym symbols (spec layout: nsym cumulative values + a counter)
void cdf_bits(const uint16_t *cdf, int nsym, int nc, float *out) {
    for (int k = 0; k < nc; k++, cdf += nsym + 1, out += nsym)
        for (int s = 0; s < nsym; s++) {
            const double p = std::max(((double)cdf[s] - (s ? (double)cdf[s - 1] : 0.0)) / 32768.0, 1.0 / 32768.0);
            out[s] = (float)-std::log2(p);
        }
}
template <typename D, typename S> void fill(D &dst, const S &src, int nsym) {
    cdf_bits((const uint16_t *)&src, nsym, (int)(sizeof(D) / sizeof(float)) / nsym, (float *)&dst);
}

}  // namespace

void build_tables(DevTables &T) {
    memset(&T, 0, sizeof T);
    double n2[3][7] = {};
    int at = 0;
    for (int kind = 0; kind < 3; kind++)
        for (int ln = 2; ln <= 6; ln++) {
            if ((kind == K_ADST && ln > 4) || (kind == K_IDT && ln > 5)) continue;
            const int n = 1 << ln;
            const double A = 16384.0;
            std::vector<double> M((size_t)n * n);
            double acc = 0;
            for (int j = 0; j < n; j++) {   // impulse responses of the exact inverse transform (wide clamp range)
                int32_t t[64] = {0};
                t[j] = 16384;
                uint32_t err = 0;
                inv_1d(t, kind, ln, 31, err);
                for (int i = 0; i < n; i++) {
                    M[(size_t)i * n + j] = (double)t[i] / A;
                    acc += M[(size_t)i * n + j] * M[(size_t)i * n + j];
                }
            }
            n2[kind][ln] = acc / n;
            const std::vector<double> F = gj_inverse(M, n);
            T.fwdOff[kind][ln] = at;
            for (int i = 0; i < n; i++)
                for (int k = 0; k < n; k++) {
                    T.fwd[at + i * n + k] = (float)F[(size_t)i * n + k];
                    T.fwdT[at + k * n + i] = (float)F[(size_t)i * n + k];
                }
            at += n * n;
        }
    for (int t = 0; t < 19; t++) {
        const int lw = av1t::Tx_Width_Log2[t], lh = av1t::Tx_Height_Log2[t];
        const double rect = (lw - lh == 1 || lh - lw == 1) ? 2896.0 / 4096.0 : 1.0;
        const double s = (double)(1 << (av1t::Transform_Row_Shift[t] + 4));
        T.txScale[t] = (float)(s / rect);
        for (int ty = 0; ty < 16; ty++) {
            const double nr = n2[row_kind(ty)][lw], nc = n2[col_kind(ty)][lh];
            T.txg[t][ty] = (nr == 0 || nc == 0) ? 0.f : (float)(rect * rect * nr * nc / (s * s));
        }
    }
    for (int q = 0; q < 4; q++) {
        CostTab &C = T.cost[q];
        using namespace av1t;
        fill(C.y_mode, Default_Intra_Frame_Y_Mode_Cdf, 13);
        fill(C.uv_cfl_no, Default_Uv_Mode_Cfl_Not_Allowed_Cdf, 13);
        fill(C.uv_cfl_yes, Default_Uv_Mode_Cfl_Allowed_Cdf, 14);
        fill(C.angle_delta, Default_Angle_Delta_Cdf, 7);
        fill(C.skip, Default_Skip_Cdf, 2);
        fill(C.tx_set1, Default_Intra_Tx_Type_Set1_Cdf, 7);
        fill(C.tx_set2, Default_Intra_Tx_Type_Set2_Cdf, 5);
        fill(C.cfl_sign, Default_Cfl_Sign_Cdf, 8);
        fill(C.cfl_alpha, Default_Cfl_Alpha_Cdf, 16);
        fill(C.txb_skip, Default_Txb_Skip_Cdf[q], 2);
        fill(C.eob_pt16, Default_Eob_Pt_16_Cdf[q], 5);
        fill(C.eob_pt32, Default_Eob_Pt_32_Cdf[q], 6);
        fill(C.eob_pt64, Default_Eob_Pt_64_Cdf[q], 7);
        fill(C.eob_pt128, Default_Eob_Pt_128_Cdf[q], 8);
        fill(C.eob_pt256, Default_Eob_Pt_256_Cdf[q], 9);
        fill(C.eob_pt512, Default_Eob_Pt_512_Cdf[q], 10);
        fill(C.eob_pt1024, Default_Eob_Pt_1024_Cdf[q], 11);
        fill(C.eob_extra, Default_Eob_Extra_Cdf[q], 2);
        fill(C.dc_sign, Default_Dc_Sign_Cdf[q], 2);
        fill(C.coeff_base_eob, Default_Coeff_Base_Eob_Cdf[q], 3);
        fill(C.coeff_base, Default_Coeff_Base_Cdf[q], 4);
        fill(C.coeff_br, Default_Coeff_Br_Cdf[q], 4);
        fill(C.tx8, Default_Tx_8x8_Cdf, 2);
        fill(C.tx16, Default_Tx_16x16_Cdf, 3);
        fill(C.tx32, Default_Tx_32x32_Cdf, 3);
        fill(C.tx64, Default_Tx_64x64_Cdf, 3);
    }
    // open-loop averages over the neighbour contexts (q-context independent CDFs)
    for (int m = 0; m < 13; m++) {
        double p = 0;
        for (int a = 0; a < 5; a++)
            for (int l = 0; l < 5; l++) {
                const uint16_t *cd = av1t::Default_Intra_Frame_Y_Mode_Cdf[a][l];
                p += (cd[m] - (m ? cd[m - 1] : 0)) / 32768.0;
            }
        T.avgYMode[m] = (float)-std::log2(std::max(p / 25.0, 1e-6));
    }
    for (int b = 0; b < 5; b++)
        for (int p = 0; p < 10; p++) T.avgPart[b][p] = 20.f;
    for (int bsl = 1; bsl <= 4; bsl++)
        for (int p = 0; p < (bsl == 1 ? 4 : 10); p++) {
            double pr = 0;
            for (int ctx = 0; ctx < 4; ctx++) {
                const uint16_t *cd = bsl == 1 ? av1t::Default_Partition_W8_Cdf[ctx] : bsl == 2 ? av1t::Default_Partition_W16_Cdf[ctx]
                                   : bsl == 3 ? av1t::Default_Partition_W32_Cdf[ctx] : av1t::Default_Partition_W64_Cdf[ctx];
                pr += (cd[p] - (p ? cd[p - 1] : 0)) / 32768.0;
            }
            T.avgPart[bsl][p] = (float)-std::log2(std::max(pr / 4.0, 1e-6));
        }
    double ps = 0;
    for (int ctx = 0; ctx < 3; ctx++) ps += av1t::Default_Skip_Cdf[ctx][0] / 32768.0;
    T.avgSkip0 = (float)-std::log2(ps / 3.0);
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 256; q++) {
            const int dq = av1t::Dc_Qlookup[b][q];
            double rd = (double)dq * dq * (3.3 + 0.0015 * dq);
            int w = ((255 - q) * 3) / 4;
            w = w < 0 ? 0 : w > 72 ? 72 : w;
            rd = rd * (128 + w) / 128.0;
            T.lambda[b][q] = (float)(rd / 2048.0 * 1.0);
            T.lamS[b][q] = sqrtf(T.lambda[b][q]);
        }
    for (int t = 0; t < 19; t++) {
        if (av1t::Tx_Width[t] == 64 || av1t::Tx_Height[t] == 64) continue;
        const uint16_t *s = scan_of(t, DCT_DCT);
        const int n = av1t::Tx_Width[t] * av1t::Tx_Height[t];
        for (int k = 0; k < n; k++) T.iscan[dscan_off(t) + s[k]] = (uint16_t)k;
    }
}

}  // namespace av1m
