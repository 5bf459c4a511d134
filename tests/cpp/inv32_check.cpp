// Host check of the GPU inverse transforms' 32-bit arithmetic (av1gpu_amd/csrc/av1m_common.h, Inv) against the
// oracle's 64-bit restatement (oracle/orc_dsp.h, inv_1d): every 1-D kind and size (DCT 4..64, ADST 4..16, identity
// 4..32) at the row and column ranges of 8- and 10-bit (r = 16, 18), on random inputs inside r bits of several
// magnitudes (so that some passes violate the ranges).  The range flag must be raised on the same inputs, and where it
// is not the outputs must be equal.  Prints "ok <cases> <flagged>" or the first disagreement and exits 1.
#include <cstdio>
#include <cstdint>
#include <initializer_list>
#include "../../av1gpu_amd/csrc/av1m_common.h"
#include "../../oracle/orc_dsp.h"

static uint64_t rng = 88172645463325252ull;
static uint32_t rnd() { rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17; return (uint32_t)(rng >> 16); }

int main() {
    long cases = 0, flagged = 0;
    const int kinds[3] = {av1m::K_DCT, av1m::K_ADST, av1m::K_IDT};
    for (int ki = 0; ki < 3; ki++)
        for (int n = 2; n <= 6; n++) {
            if (kinds[ki] == av1m::K_ADST && n > 4) continue;
            if (kinds[ki] == av1m::K_IDT && n > 5) continue;
            for (int r : {16, 18})
                for (int mag = 1; mag <= r - 1; mag++)
                    for (int rep = 0; rep < 2000; rep++) {
                        int32_t a[64], b[64];
                        const int N = 1 << n;
                        const int sparse = rep & 1;
                        for (int i = 0; i < N; i++) {
                            int32_t v = (int32_t)(rnd() % (1u << mag)) - (int32_t)(1u << (mag - 1));
                            if (sparse && (rnd() & 3)) v = 0;
                            a[i] = b[i] = v;
                        }
                        uint32_t ea = 0, eb = 0;
                        switch (n) {
                        case 2: av1m::Inv::run<2>(a, kinds[ki], r, ea); break;
                        case 3: av1m::Inv::run<3>(a, kinds[ki], r, ea); break;
                        case 4: av1m::Inv::run<4>(a, kinds[ki], r, ea); break;
                        case 5: av1m::Inv::run<5>(a, kinds[ki], r, ea); break;
                        default: av1m::Inv::run<6>(a, kinds[ki], r, ea); break;
                        }
                        orc::inv_1d(b, kinds[ki], n, r, &eb);
                        cases++;
                        if ((ea != 0) != (eb != 0)) {
                            printf("flag mismatch kind %d n %d r %d mag %d: %u vs %u\n", kinds[ki], n, r, mag, ea, eb);
                            return 1;
                        }
                        if (eb) { flagged++; continue; }
                        for (int i = 0; i < N; i++)
                            if (a[i] != b[i]) {
                                printf("value mismatch kind %d n %d r %d mag %d i %d: %d vs %d\n", kinds[ki], n, r, mag, i, a[i], b[i]);
                                return 1;
                            }
                    }
        }
    // the row pre-scaling of 2:1 blocks: round2(T * 2896, 12) for every 18-bit input
    for (int32_t v = -(1 << 17); v < (1 << 17); v++) {
        cases++;
        if (av1m::Inv::rnd12(av1m::Inv::mul(v, 2896)) != (int32_t)orc::round2_64((int64_t)v * 2896, 12)) {
            printf("rect scaling mismatch at %d\n", v);
            return 1;
        }
    }
    printf("ok %ld %ld\n", cases, flagged);
    return 0;
}
