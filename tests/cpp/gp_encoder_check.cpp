// GPU check of the reference-shaped C++ API (include/av1mi_gp.hpp, av1gpu_amd/csrc/gp_encoder.cpp) against the C ABI:
// GpEncoder::encode gives the records of av1mi_encode, frame parameters as the reference sets them, the reconstruction
// cropped to the visible size; with GpDevRecords the same records stay on the device; gp_frames_to_10bit maps every
// sample as round(v * 1023 / 255) (chroma around 512); unimplemented options are refused.  Prints "gp ok", exits 0.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <cmath>
#include <string>
#include <vector>
#include "av1mi_gp.hpp"

static int fails = 0;
#define CHECK(c) do { if (!(c)) { printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); fflush(stdout); fails++; } } while (0)

struct Planes {   // a padded 8-bit 4:2:0 synthetic frame in device memory
    int W, H, aw, ah;
    uint8_t *d[3] = {nullptr, nullptr, nullptr};
    std::vector<uint8_t> h[3];
    Planes(uint64_t seed, int W_, int H_) : W(W_), H(H_) {
        aw = (W + 63) / 64 * 64; ah = (H + 63) / 64 * 64;
        const int cw = (W + 1) / 2, ch = (H + 1) / 2;
        std::vector<uint8_t> y((size_t)W * H), u((size_t)cw * ch), v((size_t)cw * ch);
        const int st[3] = {W, cw, cw};
        av1mi_synth_image(seed, W, H, 1, 8, y.data(), u.data(), v.data(), st);
        const std::vector<uint8_t> *src[3] = {&y, &u, &v};
        for (int p = 0; p < 3; p++) {
            const int w = p ? cw : W, hh = p ? ch : H, pw = p ? aw / 2 : aw, ph = p ? ah / 2 : ah;
            h[p].resize((size_t)pw * ph);
            for (int r = 0; r < ph; r++)
                for (int c = 0; c < pw; c++) h[p][(size_t)r * pw + c] = (*src[p])[(size_t)std::min(r, hh - 1) * w + std::min(c, w - 1)];
            hipMalloc((void **)&d[p], h[p].size());
            hipMemcpy(d[p], h[p].data(), h[p].size(), hipMemcpyHostToDevice);
        }
    }
    ~Planes() { for (auto p : d) hipFree(p); }
    av1mi::GpFrameIn in(int q) const {
        av1mi::GpFrameIn f;
        f.W = W; f.H = H; f.aw = aw; f.ah = ah; f.qindex = q;
        for (int p = 0; p < 3; p++) { f.dsrc[p] = d[p]; f.spitch[p] = p ? aw / 2 : aw; }
        return f;
    }
};

int main() {
    std::string err;
    av1mi::GpConfig cfg;
    cfg.qindex = 120; cfg.edgeFilter = 1;
    cfg.P.K = 6; cfg.P.Krecheck = 6; cfg.P.txAllCL = 1; cfg.P.uvK = 6;
    av1mi::GpEncoder enc;
    CHECK(enc.init(cfg, err) == 0);
    Planes a(1000, 200, 136), b(1001, 96, 70);
    std::vector<av1mi::GpFrameIn> in = {a.in(120), b.in(80)};
    in[1].colorMatrix = 6;
    std::vector<av1mi::GpFrameOut> out;
    av1mi::GpTimes t;
    CHECK(enc.encode(in, out, t, err) == 0);
    CHECK(out.size() == 2 && t.sbs == 4 * 3 + 2 * 2 && t.blocks > 0);
    // the same frames through the C ABI
    av1mi_config c;
    memset(&c, 0, sizeof c);
    av1mi_default_params(&c.P);
    c.qindex = 120; c.edgeFilter = 1;
    av1mi_encoder *h = nullptr;
    char e[256];
    CHECK(av1mi_create(&c, &h, e, sizeof e) == 0);
    av1mi_frame_in fi[2];
    memset(fi, 0, sizeof fi);
    const Planes *pl[2] = {&a, &b};
    for (int i = 0; i < 2; i++) {
        fi[i].W = pl[i]->W; fi[i].H = pl[i]->H; fi[i].aw = pl[i]->aw; fi[i].ah = pl[i]->ah; fi[i].bitDepth = 8;
        fi[i].qindex = i ? 80 : 120;
        for (int p = 0; p < 3; p++) { fi[i].dsrc[p] = pl[i]->d[p]; fi[i].spitch[p] = p ? pl[i]->aw / 2 : pl[i]->aw; }
    }
    av1mi_frame_out fo[2];
    av1mi_times tm;
    CHECK(av1mi_encode(h, fi, 2, fo, &tm, nullptr, 1, e, sizeof e) == 0);
    for (int i = 0; i < 2; i++) {
        const av1mi::GpFrameOut &g = out[i];
        CHECK(g.blocks.size() == fo[i].n_blocks && !memcmp(g.blocks.data(), fo[i].blocks, fo[i].n_blocks * 16));
        CHECK(g.block_txb.size() == fo[i].n_blocks + 1 && !memcmp(g.block_txb.data(), fo[i].block_txb, (fo[i].n_blocks + 1) * 4));
        CHECK(g.txb.size() == fo[i].n_txb && !memcmp(g.txb.data(), fo[i].txb, fo[i].n_txb * 16));
        CHECK(g.coeffs.size() == fo[i].n_coeffs && !memcmp(g.coeffs.data(), fo[i].coeffs, fo[i].n_coeffs * 2));
        CHECK(g.fp.version == 5 && g.fp.width == (uint32_t)pl[i]->W && g.fp.height == (uint32_t)pl[i]->H);
        CHECK(g.fp.base_q_idx == (i ? 80 : 120) && g.fp.bit_depth == 8 && g.fp.chroma_format == 0 && g.fp.enable_intra_edge_filter == 1);
        CHECK(g.fp.tile_cols_log2 == -1 && g.fp.tile_rows_log2 == -1 && g.fp.tx_mode_select == 0);
        CHECK(i ? (g.fp.color_description == 1 && g.fp.matrix_coefficients == 6 && g.fp.color_primaries == 1 && g.fp.transfer_characteristics == 13)
                : g.fp.color_description == 0);
        for (int p = 0; p < 3; p++) {   // cropped reconstruction = the visible part of the padded one
            const int w = p ? (pl[i]->W + 1) / 2 : pl[i]->W, hh = p ? (pl[i]->H + 1) / 2 : pl[i]->H;
            CHECK(g.recon[p].size() == (size_t)w * hh);
            bool same = true;
            for (int r = 0; r < hh && same; r++)
                same = !memcmp(g.recon[p].data() + (size_t)r * w, (const uint8_t *)fo[i].recon[p] + (size_t)r * fo[i].recon_pitch[p], w);
            CHECK(same);
        }
        // record() feeds the library's writer
        av1mi_frame_out rec = out[i].record();
        void *data = nullptr;
        size_t size = 0;
        CHECK(av1mi_write_avif(&rec, 1, &data, &size, e, sizeof e) == 0 && size > 100);
        av1mi_buffer_free(data);
    }
    // device-resident records: the same bytes, frame-relative block_txb / coeff_off, frames back to back
    av1mi::GpDevRecords dev;
    std::vector<av1mi::GpFrameOut> out2;
    CHECK(enc.encode(in, out2, t, err, 0, &dev) == 0);
    CHECK(dev.blk0.size() == 3 && dev.blk0[1] == fo[0].n_blocks && dev.blk0[2] == fo[0].n_blocks + fo[1].n_blocks);
    CHECK(out2[0].blocks.empty() && out2[1].fp.base_q_idx == 80);
    for (int i = 0; i < 2; i++) {
        std::vector<av1mi_block> hb(fo[i].n_blocks);
        std::vector<uint32_t> hbt(fo[i].n_blocks);
        std::vector<av1mi_txb> ht(fo[i].n_txb);
        std::vector<int16_t> hc(fo[i].n_coeffs);
        hipMemcpy(hb.data(), dev.blocks + dev.blk0[i], hb.size() * 16, hipMemcpyDeviceToHost);
        hipMemcpy(hbt.data(), dev.blockTxbRel + dev.blk0[i], hbt.size() * 4, hipMemcpyDeviceToHost);
        hipMemcpy(ht.data(), dev.txb + dev.txb0[i], ht.size() * 16, hipMemcpyDeviceToHost);
        hipMemcpy(hc.data(), dev.coeffs + dev.coef0[i], hc.size() * 2, hipMemcpyDeviceToHost);
        CHECK(!memcmp(hb.data(), fo[i].blocks, hb.size() * 16) && !memcmp(hbt.data(), fo[i].block_txb, hbt.size() * 4));
        CHECK(!memcmp(ht.data(), fo[i].txb, ht.size() * 16) && !memcmp(hc.data(), fo[i].coeffs, hc.size() * 2));
        CHECK(!memcmp(&dev.fp[i], &out[i].fp, sizeof(av1mi_frame_params)));
    }
    av1mi::gp_dev_records_free(dev, 0);
    hipDeviceSynchronize();
    CHECK(dev.pool == nullptr);
    for (int i = 0; i < 2; i++) av1mi_frame_out_free(&fo[i]);
    av1mi_destroy(h);
    // gp_frames_to_10bit: every sample of every plane
    std::vector<av1mi::GpFrameIn> f10 = {a.in(120)};
    uint8_t *buf = nullptr;
    CHECK(av1mi::gp_frames_to_10bit(f10, 0, &buf, err) == 0 && buf && f10[0].bitDepth == 10);
    hipDeviceSynchronize();
    for (int p = 0; p < 3; p++) {
        const int pw = p ? a.aw / 2 : a.aw, ph = p ? a.ah / 2 : a.ah;
        CHECK(f10[0].spitch[p] == pw);
        std::vector<uint16_t> o((size_t)pw * ph);
        hipMemcpy(o.data(), f10[0].dsrc[p], o.size() * 2, hipMemcpyDeviceToHost);
        int bad = 0;
        for (size_t k = 0; k < o.size(); k++) {
            const double v = a.h[p][k];
            const long want = p ? std::lround((v - 128.0) * 1023.0 / 255.0) + 512 : std::lround(v * 1023.0 / 255.0);
            bad += o[k] != (uint16_t)std::min(1023L, std::max(0L, want));
        }
        CHECK(bad == 0);
    }
    // the 10-bit frame encodes
    std::vector<av1mi::GpFrameOut> out10;
    CHECK(enc.encode(f10, out10, t, err) == 0 && out10[0].fp.bit_depth == 10 && out10[0].recon[0].size() == (size_t)a.W * a.H * 2);
    hipFree(buf);
    // options outside this core are refused with their names
    av1mi::GpConfig bad = cfg;
    bad.qm = 5;
    av1mi::GpEncoder e2;
    CHECK(e2.init(bad, err) == -1 && err.find("qm") != std::string::npos);
    bad = cfg;
    bad.P.rdoq = 1;
    CHECK(e2.init(bad, err) == -1 && err.find("rdoq") != std::string::npos);
    std::vector<av1mi::GpFrameIn> withQ = {a.in(120)};
    withQ[0].sbQ.assign(12, 100);
    CHECK(enc.encode(withQ, out, t, err) == -1 && err.find("sbQ") != std::string::npos);
    av1mi::GpParams none;
    none.shapeMask = 1 << 12;   // 64x64 blocks only: frames whose coded area ends inside a superblock have no partition
    CHECK(av1mi::gp_check_partitions(none, err) == -1);
    CHECK(av1mi::gp_check_partitions(av1mi::GpParams(), err) == 0);
    if (fails) { printf("%d failures\n", fails); return 1; }
    printf("gp ok\n");
    return 0;
}
