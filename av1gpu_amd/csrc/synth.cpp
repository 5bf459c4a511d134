// Seeded synthetic test images (the benchmark workloads' generator).
//   luma   pink noise (1/f amplitude: an octave sum of bilinear value noise whose amplitude doubles with the grid
//          spacing), normalised to mean 128 and std 40, plus 20 oriented step edges (contrast U[20, 120]) and Gaussian
//          noise of sigma 2, rounded and clipped to [0, 255] (each edge a straight step inside a disk of radius
//          U[0.05, 0.30] x the longer side around a random point);
//   chroma low-pass value noise (coarse octaves only) at the chroma resolution, std 12 around 128;
//   10-bit images scale every amplitude by 1023 / 255 and clip to [0, 1023].
// Deterministic: splitmix64 streams and double arithmetic in a fixed order, so the same seed gives the same planes on
// every machine with the same C library.
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
    double gauss() {   // Box-Muller, one value per call
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

// octave sum of bilinear value noise, spacings 2^o0 .. 2^o1 pixels, amplitude proportional to the spacing
void value_noise(Rng &r, int w, int h, int o0, int o1, std::vector<double> &out) {
    out.assign((size_t)w * h, 0.0);
    for (int o = o0; o <= o1; o++) {
        const int s = 1 << o;
        const int gw = w / s + 2, gh = h / s + 2;
        std::vector<double> g((size_t)gw * gh);
        for (auto &v : g) v = r.uniform() * 2.0 - 1.0;
        const double amp = (double)s, inv = 1.0 / s;
        for (int y = 0; y < h; y++) {
            const int gy = y >> o;
            const double fy = (y - (gy << o)) * inv;
            for (int x = 0; x < w; x++) {
                const int gx = x >> o;
                const double fx = (x - (gx << o)) * inv;
                const double a = g[(size_t)gy * gw + gx], b = g[(size_t)gy * gw + gx + 1];
                const double c = g[(size_t)(gy + 1) * gw + gx], d = g[(size_t)(gy + 1) * gw + gx + 1];
                const double top = a + (b - a) * fx, bot = c + (d - c) * fx;
                out[(size_t)y * w + x] += amp * (top + (bot - top) * fy);
            }
        }
    }
}

void normalise(std::vector<double> &v, double mean, double sd) {
    double s = 0, s2 = 0;
    for (double x : v) s += x;
    const double m = s / (double)v.size();
    for (double x : v) s2 += (x - m) * (x - m);
    const double k = sd / sqrt(s2 / (double)v.size() + 1e-30);
    for (double &x : v) x = mean + (x - m) * k;
}

template <typename P>
void store(const std::vector<double> &v, int w, int h, P *dst, int stride, int maxv) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double f = floor(v[(size_t)y * w + x] + 0.5);
            f = f < 0 ? 0 : f > maxv ? maxv : f;
            dst[(size_t)y * stride + x] = (P)f;
        }
}

int top_octave(int w, int h) {
    int o = 0;
    while ((2 << o) < (w > h ? w : h) && o < 9) o++;
    return o;
}

}  // namespace

extern "C" {

// Fill the planes of a W x H image (chroma ((W + ssx) >> ssx) x ((H + ssx) >> ssx); ssx 1 = 4:2:0, 0 = 4:4:4) from seed.
// bd 8: uint8_t samples, bd 10: uint16_t samples; strides in samples.  Returns 0, or -1 for bad arguments.
int av1mi_synth_image(uint64_t seed, int W, int H, int ssx, int bd, void *y, void *u, void *v, const int *strides) {
    if (W <= 0 || H <= 0 || (bd != 8 && bd != 10) || (ssx != 0 && ssx != 1) || !y || !u || !v || !strides) return -1;
    const double sc = bd == 10 ? 1023.0 / 255.0 : 1.0;
    const int maxv = (1 << bd) - 1;
    Rng r{seed * 0x2545F4914F6CDD1Dull + 1};
    std::vector<double> L;
    value_noise(r, W, H, 0, top_octave(W, H), L);
    normalise(L, 128.0 * sc, 40.0 * sc);
    for (int e = 0; e < 20; e++) {
        // a straight step through (px, py) at angle th, limited to the disk of radius rad around that point
        const double px = r.uniform() * W, py = r.uniform() * H, th = r.uniform() * 3.141592653589793;
        const double c = (20.0 + 100.0 * r.uniform()) * sc * 0.5, cs = cos(th), sn = sin(th);
        const double rad = (0.05 + 0.25 * r.uniform()) * (W > H ? W : H), rad2 = rad * rad;
        const int y0 = (int)fmax(0.0, floor(py - rad)), y1 = (int)fmin((double)H, ceil(py + rad));
        const int x0 = (int)fmax(0.0, floor(px - rad)), x1 = (int)fmin((double)W, ceil(px + rad));
        for (int yy = y0; yy < y1; yy++)
            for (int xx = x0; xx < x1; xx++) {
                const double dx = xx - px, dy = yy - py;
                if (dx * dx + dy * dy <= rad2) L[(size_t)yy * W + xx] += (dx * cs + dy * sn >= 0.0) ? c : -c;
            }
    }
    for (double &x : L) x += 2.0 * sc * r.gauss();
    if (bd == 8) store(L, W, H, (uint8_t *)y, strides[0], maxv);
    else store(L, W, H, (uint16_t *)y, strides[0], maxv);
    const int cw = (W + ssx) >> ssx, ch = (H + ssx) >> ssx;
    void *cp[2] = {u, v};
    for (int p = 0; p < 2; p++) {
        std::vector<double> C;
        const int top = top_octave(cw, ch);
        value_noise(r, cw, ch, top > 3 - ssx ? 3 - ssx : top, top, C);
        normalise(C, 128.0 * sc, 12.0 * sc);
        if (bd == 8) store(C, cw, ch, (uint8_t *)cp[p], strides[1 + p], maxv);
        else store(C, cw, ch, (uint16_t *)cp[p], strides[1 + p], maxv);
    }
    return 0;
}

}  // extern "C"
