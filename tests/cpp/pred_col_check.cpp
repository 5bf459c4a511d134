// Host check of pred_col (av1gpu_amd/csrc/av1m_common.h) against pred_px, sample by sample: every mode (DC with its
// value), every angle delta, every block size of 4..64 on each side, without and with upsampled edges, runs of 1, 4 and
// 8 rows at strides 1 and 2 from every start that stays in the block, every column, on random 8- and 10-bit edges.
// Prints "ok <cases>" or the first mismatch and exits 1.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include "../../av1gpu_amd/csrc/av1m_common.h"

using namespace av1m;

static uint32_t rng = 12345;
static int rnd(int n) { rng = rng * 1664525u + 1013904223u; return (int)((rng >> 8) % (uint32_t)n); }

template <int CH>
static bool run(int mode, int pa, int dcv, int lw, int lh, const uint16_t *A, const uint16_t *L, int upA, int upL, long &cases) {
    const int w = 1 << lw, h = 1 << lh;
    for (int di = 1; di <= 2; di++)
        for (int i0 = 0; i0 + (CH - 1) * di < h; i0++)
            for (int j = 0; j < w; j++) {
                int o[CH];
                pred_col<CH>(mode, pa, dcv, i0, di, j, lw, lh, A, L, upA, upL, o);
                for (int c = 0; c < CH; c++) {
                    const int i = i0 + c * di;
                    const int want = mode == DC_PRED ? dcv : pred_px(mode, pa, i, j, lw, lh, A, L, upA, upL);
                    cases++;
                    if (o[c] != want) {
                        printf("mismatch mode %d pa %d %dx%d up %d/%d row %d col %d: %d != %d\n", mode, pa, w, h, upA, upL, i, j, o[c], want);
                        return false;
                    }
                }
            }
    return true;
}

int main() {
    static uint16_t abuf[512], lbuf[512];
    long cases = 0;
    for (int bd = 8; bd <= 10; bd += 2)
        for (int lw = 2; lw <= 6; lw++)
            for (int lh = 2; lh <= 6; lh++)
                for (int up = 0; up < 4; up++) {
                    for (int k = 0; k < 512; k++) { abuf[k] = (uint16_t)rnd(1 << bd); lbuf[k] = (uint16_t)rnd(1 << bd); }
                    const uint16_t *A = abuf + 64, *L = lbuf + 64;
                    const int upA = up & 1, upL = up >> 1, dcv = rnd(1 << bd);
                    for (int mode = 0; mode <= PAETH_PRED; mode++)
                        for (int d = -3; d <= 3; d++) {
                            if (!is_dir(mode) && d) continue;
                            const int pa = pangle(mode, d);
                            if (!run<1>(mode, pa, dcv, lw, lh, A, L, upA, upL, cases) || !run<4>(mode, pa, dcv, lw, lh, A, L, upA, upL, cases) ||
                                !run<8>(mode, pa, dcv, lw, lh, A, L, upA, upL, cases))
                                return 1;
                        }
                }
    printf("ok %ld\n", cases);
    return 0;
}
