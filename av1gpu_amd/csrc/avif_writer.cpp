// Host AV1 bitstream writer and AVIF muxer for a DecisionRecord (the intra decisions, tx blocks and quantized levels
// av1mi_encode returns): sequence header, key-frame header and tile data of one still picture, entropy coded with the
// CDFs adapting from the defaults within each tile, as the reference's writer codes, wrapped in an AVIF (HEIF) file.  The reference's writer is av1core/src/encoder.cpp:1072-1659 (headers 1317-1441, tiles 1290-1315,
// blocks 491-723, coefficients 1073-1288) with avif.cpp; this one follows the AV1 specification's syntax (sections
// 5.5, 5.9, 5.11) directly and codes symbols with its own range encoder, the inverse of the specification's symbol
// decoder (8.2.6).
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>
#include "../../include/av1mi.h"
#include "av1m_common.h"

namespace {
using namespace av1m;

struct WriteError : std::runtime_error {
    explicit WriteError(const std::string &m) : std::runtime_error(m) {}
};

// ---- MSB-first bit writer for the headers ----
struct Bits {
    std::vector<uint8_t> b;
    int used = 0;   // bits used in the last byte
    void bit(int x) {
        if (used == 0) b.push_back(0);
        if (x) b.back() |= (uint8_t)(0x80 >> used);
        used = (used + 1) & 7;
    }
    void put(uint32_t v, int n) {
        for (int i = n - 1; i >= 0; i--) bit((v >> i) & 1);
    }
    void trailing() {   // trailing_bits(): a one, then zeros to the byte boundary
        bit(1);
        while (used) bit(0);
    }
    void align() {
        while (used) bit(0);
    }
};

// ---- range encoder: the inverse of the specification's symbol decoder (8.2.6) ----
// The code value is a binary fraction.  The current interval is [E * 2^nb + low, + rng) in units of 2^-(emitted bits
// + nb), E being the bytes already emitted (which a carry may still increment).  A symbol s of an N-symbol CDF
// narrows it to the decoder's [cur(s), cur(s - 1)) counted from the top of the interval (the decoder inverts the code
// bits), with cur(-1) = rng.
struct RangeEnc {
    std::vector<uint8_t> out;
    uint64_t low = 0;
    uint32_t rng = 0x8000;
    int nb = 15;
    static uint32_t cur(uint32_t r, const uint16_t *cdf, int s, int N) {
        return ((r >> 8) * ((uint32_t)(32768 - cdf[s]) >> 6) >> 1) + 4u * (uint32_t)(N - s - 1);
    }
    void carry() {
        for (size_t i = out.size(); i-- > 0;) {
            if (++out[i] != 0) return;
        }
        throw WriteError("range coder carry out of the code");
    }
    void symbol(int s, const uint16_t *cdf, int N) {
        if (s < 0 || s >= N) throw WriteError("symbol out of its alphabet");
        const uint32_t u = s == 0 ? rng : cur(rng, cdf, s - 1, N), v = cur(rng, cdf, s, N);
        if (!(u > v)) throw WriteError("empty symbol interval");
        low += rng - u;
        rng = u - v;
        int d = 0;
        while ((rng << d) < 0x8000) d++;
        rng <<= d;
        low <<= d;
        nb += d;
        if (low >> nb) { carry(); low &= ((uint64_t)1 << nb) - 1; }
        while (nb >= 40) {
            out.push_back((uint8_t)(low >> (nb - 8)));
            nb -= 8;
            low &= ((uint64_t)1 << nb) - 1;
        }
    }
    // symbol s with the CDF's adaptation after it (8.2.6, disable_cdf_update = 0); cdf[N] counts the updates
    void adapt(int s, uint16_t *cdf, int N) {
        symbol(s, cdf, N);
        const int rate = 3 + (cdf[N] > 15) + (cdf[N] > 31) + (N >= 4 ? 2 : N >= 2 ? 1 : 0);
        uint32_t tmp = 0;
        for (int i = 0; i < N - 1; i++) {
            if (i == s) tmp = 1u << 15;
            if (tmp < cdf[i]) cdf[i] -= (uint16_t)((cdf[i] - tmp) >> rate);
            else cdf[i] += (uint16_t)((tmp - cdf[i]) >> rate);
        }
        cdf[N] += cdf[N] < 32;
    }
    void bit(int b) {
        static const uint16_t half[2] = {16384, 32768};
        symbol(b ? 1 : 0, half, 2);
    }
    void literal(uint32_t v, int n) {
        for (int i = n - 1; i >= 0; i--) bit((v >> i) & 1);
    }
    // the final code value: the interval's first multiple of 2^14 with bit 14 set (a one, then zero padding)
    std::vector<uint8_t> finish() {
        uint64_t e = ((low + 0x3FFF) & ~(uint64_t)0x3FFF) | 0x4000;
        if (e >> nb) { carry(); e &= ((uint64_t)1 << nb) - 1; }
        int bits = nb - 14;   // bits down to and including the marker
        int sh = nb;
        while (bits > 0) {
            sh -= 8;
            out.push_back((uint8_t)(sh >= 0 ? e >> sh : e << -sh));
            bits -= 8;
        }
        return out;
    }
};

struct Tiles {
    int colsLog2 = 0, rowsLog2 = 0, minLog2Cols = 0, maxLog2Cols = 0, minLog2Rows = 0, maxLog2Rows = 0;
    std::vector<int> colStart, rowStart;   // mi units, with the end
};
int tile_log2(int blk, int target) {
    int k = 0;
    while ((blk << k) < target) k++;
    return k;
}
// tile_info() with uniform spacing and the fewest tiles the specification allows (the layout av1mi_encode codes)
Tiles tile_info(int MiCols, int MiRows) {
    Tiles t;
    const int sbCols = (MiCols + 15) >> 4, sbRows = (MiRows + 15) >> 4;
    const int maxTileWidthSb = 4096 >> 6, maxTileAreaSb = (4096 * 2304) >> 12;
    t.minLog2Cols = tile_log2(maxTileWidthSb, sbCols);
    t.maxLog2Cols = tile_log2(1, std::min(sbCols, 64));
    t.maxLog2Rows = tile_log2(1, std::min(sbRows, 64));
    const int minLog2Tiles = std::max(t.minLog2Cols, tile_log2(maxTileAreaSb, sbRows * sbCols));
    t.colsLog2 = t.minLog2Cols;
    const int wSb = (sbCols + (1 << t.colsLog2) - 1) >> t.colsLog2;
    for (int i = 0, s = 0; s < sbCols; i++, s += wSb) t.colStart.push_back(s << 4);
    t.colStart.push_back(MiCols);
    t.minLog2Rows = std::max(minLog2Tiles - t.colsLog2, 0);
    t.rowsLog2 = t.minLog2Rows;
    const int hSb = (sbRows + (1 << t.rowsLog2) - 1) >> t.rowsLog2;
    for (int i = 0, s = 0; s < sbRows; i++, s += hSb) t.rowStart.push_back(s << 4);
    t.rowStart.push_back(MiRows);
    return t;
}

// the adaptive CDFs of one tile (initialised from the defaults of the frame's coefficient q-context)
struct Cdfs {
    uint16_t partW8[4][5], partW16[4][11], partW32[4][11], partW64[4][11];
    uint16_t skip[3][3], yMode[5][5][14], uvCfl[13][15], uvNoCfl[13][14], angle[8][8], cflSign[9], cflAlpha[6][17];
    uint16_t set1[2][13][8], set2[3][13][6];
    uint16_t txbSkip[5][13][3], eob16[2][2][6], eob32[2][2][7], eob64[2][2][8], eob128[2][2][9], eob256[2][2][10];
    uint16_t eob512[2][11], eob1024[2][12], eobExtra[5][2][9][3], dcSign[2][3][3], baseEob[5][2][4][4];
    uint16_t base[5][2][42][5], br[5][2][21][5];
    void init(int q) {
        using namespace av1t;
        memcpy(partW8, Default_Partition_W8_Cdf, sizeof partW8); memcpy(partW16, Default_Partition_W16_Cdf, sizeof partW16);
        memcpy(partW32, Default_Partition_W32_Cdf, sizeof partW32); memcpy(partW64, Default_Partition_W64_Cdf, sizeof partW64);
        memcpy(skip, Default_Skip_Cdf, sizeof skip); memcpy(yMode, Default_Intra_Frame_Y_Mode_Cdf, sizeof yMode);
        memcpy(uvCfl, Default_Uv_Mode_Cfl_Allowed_Cdf, sizeof uvCfl); memcpy(uvNoCfl, Default_Uv_Mode_Cfl_Not_Allowed_Cdf, sizeof uvNoCfl);
        memcpy(angle, Default_Angle_Delta_Cdf, sizeof angle); memcpy(cflSign, Default_Cfl_Sign_Cdf, sizeof cflSign);
        memcpy(cflAlpha, Default_Cfl_Alpha_Cdf, sizeof cflAlpha);
        memcpy(set1, Default_Intra_Tx_Type_Set1_Cdf, sizeof set1); memcpy(set2, Default_Intra_Tx_Type_Set2_Cdf, sizeof set2);
        memcpy(txbSkip, Default_Txb_Skip_Cdf[q], sizeof txbSkip);
        memcpy(eob16, Default_Eob_Pt_16_Cdf[q], sizeof eob16); memcpy(eob32, Default_Eob_Pt_32_Cdf[q], sizeof eob32);
        memcpy(eob64, Default_Eob_Pt_64_Cdf[q], sizeof eob64); memcpy(eob128, Default_Eob_Pt_128_Cdf[q], sizeof eob128);
        memcpy(eob256, Default_Eob_Pt_256_Cdf[q], sizeof eob256); memcpy(eob512, Default_Eob_Pt_512_Cdf[q], sizeof eob512);
        memcpy(eob1024, Default_Eob_Pt_1024_Cdf[q], sizeof eob1024); memcpy(eobExtra, Default_Eob_Extra_Cdf[q], sizeof eobExtra);
        memcpy(dcSign, Default_Dc_Sign_Cdf[q], sizeof dcSign); memcpy(baseEob, Default_Coeff_Base_Eob_Cdf[q], sizeof baseEob);
        memcpy(base, Default_Coeff_Base_Cdf[q], sizeof base); memcpy(br, Default_Coeff_Br_Cdf[q], sizeof br);
    }
};
static_assert(sizeof(((Cdfs *)0)->base) == sizeof(av1t::Default_Coeff_Base_Cdf[0]), "cdf dims");
static_assert(sizeof(((Cdfs *)0)->yMode) == sizeof(av1t::Default_Intra_Frame_Y_Mode_Cdf), "cdf dims");

class Writer {
public:
    Writer(const av1mi_frame_out &r, int edgeFilter) : R(r), edge(edgeFilter) {
        W = (int)R.width; H = (int)R.height;
        MiCols = 2 * ((W + 7) >> 3); MiRows = 2 * ((H + 7) >> 3);
        mono = R.monochrome; bd = R.bit_depth == 10 ? 10 : 8;
        if (R.chroma_format > 1) throw WriteError("chroma format must be 4:2:0 (0) or 4:4:4 (1)");
        ss = (R.chroma_format == 1 && !mono) ? 0 : 1;   // 4:4:4: profile 1, chroma at full resolution
        if (R.base_q_idx < 1 || R.base_q_idx > 255) throw WriteError("base_q_idx outside 1..255");
        qctx = R.base_q_idx <= 20 ? 0 : R.base_q_idx <= 60 ? 1 : R.base_q_idx <= 120 ? 2 : 3;
        tiles = tile_info(MiCols, MiRows);
        if ((int)tiles.colStart.size() - 1 != R.tile_cols || (int)tiles.rowStart.size() - 1 != R.tile_rows)
            throw WriteError("the record's tile layout is not the minimum uniform one");
        // post-filter parameters (loop_filter_params, cdef_params) from the frame's decisions, checked as the reference's
        // av1g_validate does (av1core/src/encoder.cpp:1495-1524)
        if (R.fp.version) {
            const av1mi_frame_params &F = R.fp;
            for (int k = 0; k < 4; k++) {
                if (F.loop_filter_level[k] > 63) throw WriteError("loop_filter_level > 63");
                lf[k] = F.loop_filter_level[k];
            }
            if (mono) lf[2] = lf[3] = 0;
            if (!lf[0] && !lf[1] && (lf[2] || lf[3])) throw WriteError("chroma loop filter levels need a nonzero luma level");
            if (F.loop_filter_sharpness > 7) throw WriteError("loop_filter_sharpness > 7");
            lfSharp = F.loop_filter_sharpness;
            if (F.enable_cdef > 1 || F.cdef_damping_minus_3 > 3 || F.cdef_bits > 3) throw WriteError("bad CDEF parameters");
            cdefOn = F.enable_cdef;
            if (cdefOn) {
                cdefDamp3 = F.cdef_damping_minus_3;
                cdefBits = F.cdef_bits;
                auto secOk = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4; };
                for (int i = 0; i < (1 << cdefBits); i++) {
                    if (F.cdef_y_pri[i] > 15 || F.cdef_uv_pri[i] > 15 || !secOk(F.cdef_y_sec[i]) || !secOk(F.cdef_uv_sec[i]))
                        throw WriteError("bad CDEF strength");
                    cdefYPri[i] = F.cdef_y_pri[i]; cdefYSec[i] = F.cdef_y_sec[i];
                    cdefUVPri[i] = F.cdef_uv_pri[i]; cdefUVSec[i] = F.cdef_uv_sec[i];
                }
            }
        }
        const size_t mi = (size_t)MiCols * MiRows;
        mSize.assign(mi, 0); mSkip.assign(mi, 0); mYMode.assign(mi, 0);
        for (int p = 0; p < 3; p++) {
            aLev[p].assign(MiCols + 32, 0); aDc[p].assign(MiCols + 32, 0);
            lLev[p].assign(MiRows + 32, 0); lDc[p].assign(MiRows + 32, 0);
        }
    }

    std::vector<uint8_t> obus() {
        std::vector<uint8_t> out;
        Bits seq = sequence_header();
        obu(out, 1, seq.b);   // OBU_SEQUENCE_HEADER
        Bits fh = frame_header();
        std::vector<uint8_t> frame = fh.b;
        const int nt = (int)(tiles.colStart.size() - 1) * (int)(tiles.rowStart.size() - 1);
        // tile_group_obu: with several tiles, tile_start_and_end_present_flag = 0 then byte alignment
        if (nt > 1) frame.push_back(0);
        size_t blk = 0;
        for (size_t tr = 0; tr + 1 < tiles.rowStart.size(); tr++)
            for (size_t tc = 0; tc + 1 < tiles.colStart.size(); tc++) {
                const std::vector<uint8_t> td = tile((int)tr, (int)tc, blk);
                const bool last = tr + 2 == tiles.rowStart.size() && tc + 2 == tiles.colStart.size();
                if (!last) {   // tile_size_minus_1, le(TileSizeBytes = 4)
                    const uint32_t s = (uint32_t)td.size() - 1;
                    for (int k = 0; k < 4; k++) frame.push_back((uint8_t)(s >> (8 * k)));
                }
                frame.insert(frame.end(), td.begin(), td.end());
            }
        if (blk != R.n_blocks) throw WriteError("the record has blocks outside its partition trees");
        obu(out, 6, frame);   // OBU_FRAME
        seqObu.assign(out.begin(), out.begin() + (long)seqLen);
        return out;
    }
    std::vector<uint8_t> seqObu;   // the sequence header OBU (for the av1C box)
    int profile() const { return ss ? 0 : 1; }
    int subsampling() const { return ss; }
    int level() const { return 31; }

private:
    const av1mi_frame_out &R;
    int edge, W, H, MiCols, MiRows, mono, bd, qctx, ss = 1;
    int lf[4] = {0, 0, 0, 0}, lfSharp = 0, cdefOn = 0, cdefDamp3 = 0, cdefBits = 0;
    int cdefYPri[8] = {}, cdefYSec[8] = {}, cdefUVPri[8] = {}, cdefUVSec[8] = {};
    int sbCdef = -1;   // cdef_idx of the current superblock (-1: none read yet; spec clear_cdef / read_cdef)
    size_t seqLen = 0;
    Tiles tiles;
    std::vector<uint8_t> mSize, mSkip, mYMode;
    Cdfs cdf;
    std::vector<uint8_t> aLev[3], aDc[3], lLev[3], lDc[3];
    int tr0 = 0, tr1 = 0, tc0 = 0, tc1 = 0;

    void obu(std::vector<uint8_t> &out, int type, const std::vector<uint8_t> &payload) {
        out.push_back((uint8_t)((type << 3) | 2));   // obu_has_size_field
        uint64_t n = payload.size();
        do {
            uint8_t byte = n & 0x7F;
            n >>= 7;
            out.push_back(n ? (uint8_t)(byte | 0x80) : byte);
        } while (n);
        out.insert(out.end(), payload.begin(), payload.end());
        if (type == 1) seqLen = out.size();
    }

    // sequence_header_obu, reduced still picture header (5.5)
    Bits sequence_header() {
        Bits s;
        s.put(profile(), 3);         // seq_profile: 0 Main (8/10-bit 4:2:0, monochrome), 1 High (8/10-bit 4:4:4)
        s.put(1, 1);                 // still_picture
        s.put(1, 1);                 // reduced_still_picture_header
        s.put(level(), 5);           // seq_level_idx[0]
        int wb = 1, hb = 1;
        while ((W - 1) >> wb) wb++;
        while ((H - 1) >> hb) hb++;
        s.put(wb - 1, 4); s.put(hb - 1, 4);
        s.put(W - 1, wb); s.put(H - 1, hb);
        s.put(0, 1);                 // use_128x128_superblock
        s.put(0, 1);                 // enable_filter_intra
        s.put(edge ? 1 : 0, 1);      // enable_intra_edge_filter
        s.put(0, 1);                 // enable_superres
        s.put(cdefOn, 1);            // enable_cdef
        s.put(0, 1);                 // enable_restoration
        s.put(bd == 10, 1);          // high_bitdepth
        if (profile() != 1) s.put(mono, 1);   // mono_chrome (absent in profile 1)
        s.put(1, 1);                 // color_description_present_flag
        s.put(1, 8);                 // color_primaries BT.709
        s.put(13, 8);                // transfer_characteristics sRGB
        s.put(6, 8);                 // matrix_coefficients BT.601
        s.put(1, 1);                 // color_range: full
        if (!mono) {
            if (ss) s.put(0, 2);     // chroma_sample_position: unknown (4:2:0 only; profile 1 implies 4:4:4)
            s.put(0, 1);             // separate_uv_delta_q
        }
        s.put(0, 1);                 // film_grain_params_present
        s.trailing();
        return s;
    }

    // uncompressed_header of the key frame (5.9), then byte_alignment (OBU_FRAME)
    Bits frame_header() {
        Bits f;
        f.put(0, 1);                 // disable_cdf_update: the CDFs adapt within each tile
        f.put(0, 1);                 // allow_screen_content_tools
        f.put(0, 1);                 // render_and_frame_size_different
        f.put(1, 1);                 // uniform_tile_spacing_flag
        for (int k = tiles.minLog2Cols; k < tiles.maxLog2Cols; k++) {   // increment_tile_cols_log2
            if (k < tiles.colsLog2) f.put(1, 1);
            else { f.put(0, 1); break; }
        }
        for (int k = tiles.minLog2Rows; k < tiles.maxLog2Rows; k++) {
            if (k < tiles.rowsLog2) f.put(1, 1);
            else { f.put(0, 1); break; }
        }
        if (tiles.colsLog2 + tiles.rowsLog2 > 0) {
            f.put(0, tiles.colsLog2 + tiles.rowsLog2);   // context_update_tile_id
            f.put(3, 2);                                 // tile_size_bytes_minus_1
        }
        f.put(R.base_q_idx, 8);      // base_q_idx
        f.put(0, 1);                 // DeltaQYDc
        if (!mono) { f.put(0, 1); f.put(0, 1); }   // DeltaQUDc, DeltaQUAc
        f.put(0, 1);                 // using_qmatrix
        f.put(0, 1);                 // segmentation_enabled
        f.put(0, 1);                 // delta_q_present
        f.put(lf[0], 6); f.put(lf[1], 6);    // loop_filter_params: loop_filter_level[0], [1]
        if (!mono && (lf[0] || lf[1])) { f.put(lf[2], 6); f.put(lf[3], 6); }
        f.put(lfSharp, 3);           // loop_filter_sharpness
        f.put(0, 1);                 // loop_filter_delta_enabled
        if (cdefOn) {                // cdef_params
            f.put(cdefDamp3, 2);
            f.put(cdefBits, 2);
            for (int i = 0; i < (1 << cdefBits); i++) {
                f.put(cdefYPri[i], 4);
                f.put(cdefYSec[i] == 4 ? 3 : cdefYSec[i], 2);
                if (!mono) { f.put(cdefUVPri[i], 4); f.put(cdefUVSec[i] == 4 ? 3 : cdefUVSec[i], 2); }
            }
        }
        f.put(0, 1);                 // tx_mode_select = 0: TX_MODE_LARGEST
        f.put(0, 1);                 // reduced_tx_set
        f.align();
        return f;
    }

    uint8_t &at(std::vector<uint8_t> &m, int r, int c) { return m[(size_t)r * MiCols + c]; }
    bool inside(int r, int c) const { return c >= tc0 && c < tc1 && r >= tr0 && r < tr1; }

    std::vector<uint8_t> tile(int tr, int tc, size_t &blk) {
        tr0 = tiles.rowStart[tr]; tr1 = tiles.rowStart[tr + 1]; tc0 = tiles.colStart[tc]; tc1 = tiles.colStart[tc + 1];
        RangeEnc ec;
        cdf.init(qctx);
        for (int p = 0; p < 3; p++) { std::fill(aLev[p].begin(), aLev[p].end(), 0); std::fill(aDc[p].begin(), aDc[p].end(), 0); }
        for (int r = tr0; r < tr1; r += 16) {
            for (int p = 0; p < 3; p++) { std::fill(lLev[p].begin(), lLev[p].end(), 0); std::fill(lDc[p].begin(), lDc[p].end(), 0); }
            for (int c = tc0; c < tc1; c += 16) {
                sbCdef = -1;   // clear_cdef
                partition(ec, r, c, BLOCK_64X64, blk);
            }
        }
        return ec.finish();
    }

    const RecBlock *blockAt(size_t k) const { return k < R.n_blocks ? (const RecBlock *)&R.blocks[k] : nullptr; }
    bool isAt(size_t k, int r, int c, int bs) const {
        const RecBlock *b = blockAt(k);
        return b && b->mi_row == r && b->mi_col == c && b->bsize == bs;
    }

    // the partition of the node (r, c, bSize) that produces the record's next blocks (decode order)
    int derive(int r, int c, int bSize, int hasRows, int hasCols, size_t k) const {
        const RecBlock *b = blockAt(k);
        if (!b || b->mi_row != r || b->mi_col != c) throw WriteError("the record's blocks do not tile the frame in decode order");
        const int half = bw4(bSize) >> 1, sp = av1t::Partition_Subsize[PARTITION_SPLIT][bSize];
        const int hz = av1t::Partition_Subsize[PARTITION_HORZ][bSize], vt = av1t::Partition_Subsize[PARTITION_VERT][bSize];
        const bool ext = bSize > BLOCK_8X8 && hasRows && hasCols;   // the extended partitions need a whole node
        int p;
        if (b->bsize == bSize) p = PARTITION_NONE;
        else if (b->bsize == hz) p = ext && isAt(k + 1, r + half, c, sp) ? PARTITION_HORZ_B : PARTITION_HORZ;
        else if (b->bsize == vt) p = ext && isAt(k + 1, r, c + half, sp) ? PARTITION_VERT_B : PARTITION_VERT;
        else if (ext && b->bsize == sp && isAt(k + 1, r + half, c, sp) && isAt(k + 2, r, c + half, vt)) p = PARTITION_VERT_A;
        else if (ext && b->bsize == sp && isAt(k + 1, r, c + half, sp) && isAt(k + 2, r + half, c, hz)) p = PARTITION_HORZ_A;
        else if (ext && b->bsize == av1t::Partition_Subsize[PARTITION_HORZ_4][bSize]) p = PARTITION_HORZ_4;
        else if (ext && b->bsize == av1t::Partition_Subsize[PARTITION_VERT_4][bSize]) p = PARTITION_VERT_4;
        else p = PARTITION_SPLIT;
        const bool ok = (hasRows && hasCols) || (hasCols && (p == PARTITION_HORZ || p == PARTITION_SPLIT)) ||
                        (hasRows && (p == PARTITION_VERT || p == PARTITION_SPLIT)) || p == PARTITION_SPLIT;
        if (!ok) throw WriteError("the record's partition is not codable at a frame edge");
        return p;
    }

    uint16_t *partition_cdf(int bsl, int ctx) {
        return bsl == 1 ? cdf.partW8[ctx] : bsl == 2 ? cdf.partW16[ctx] : bsl == 3 ? cdf.partW32[ctx] : cdf.partW64[ctx];
    }

    // decode_partition() (5.11.4)
    void partition(RangeEnc &ec, int r, int c, int bSize, size_t &blk) {
        if (r >= MiRows || c >= MiCols) return;
        const int num4x4 = bw4(bSize), half = num4x4 >> 1, quarter = half >> 1;
        const int hasRows = (r + half) < MiRows, hasCols = (c + half) < MiCols;
        int p = PARTITION_NONE;
        if (bSize >= BLOCK_8X8) {
            p = derive(r, c, bSize, hasRows, hasCols, blk);
            const int bsl = av1t::Mi_Width_Log2[bSize];
            const int above = inside(r - 1, c) && av1t::Mi_Width_Log2[at(mSize, r - 1, c)] < bsl;
            const int left = inside(r, c - 1) && av1t::Mi_Height_Log2[at(mSize, r, c - 1)] < bsl;
            uint16_t *pc = partition_cdf(bsl, left * 2 + above);
            const int np = bsl == 1 ? 4 : 10;
            auto P = [&](int x) -> uint32_t { return x < np ? (uint32_t)pc[x] - (x > 0 ? pc[x - 1] : 0u) : 0u; };
            if (hasRows && hasCols) {
                ec.adapt(p, pc, np);
            } else if (hasCols) {   // split_or_horz (9.3.x): the probability mass of the partitions that split vertically
                const uint32_t psum = P(PARTITION_VERT) + P(PARTITION_SPLIT) + P(PARTITION_HORZ_A) + P(PARTITION_VERT_A) +
                                      P(PARTITION_VERT_B) + P(PARTITION_VERT_4);
                const uint16_t cdf[2] = {(uint16_t)(32768 - psum), 32768};
                ec.symbol(p == PARTITION_SPLIT, cdf, 2);
            } else if (hasRows) {   // split_or_vert
                const uint32_t psum = P(PARTITION_HORZ) + P(PARTITION_SPLIT) + P(PARTITION_HORZ_A) + P(PARTITION_HORZ_B) +
                                      P(PARTITION_VERT_A) + P(PARTITION_HORZ_4);
                const uint16_t cdf[2] = {(uint16_t)(32768 - psum), 32768};
                ec.symbol(p == PARTITION_SPLIT, cdf, 2);
            }
        }
        const int sub = av1t::Partition_Subsize[p][bSize], sp = av1t::Partition_Subsize[PARTITION_SPLIT][bSize];
        switch (p) {
        case PARTITION_NONE: block(ec, r, c, sub, blk); break;
        case PARTITION_HORZ:
            block(ec, r, c, sub, blk);
            if (hasRows) block(ec, r + half, c, sub, blk);
            break;
        case PARTITION_VERT:
            block(ec, r, c, sub, blk);
            if (hasCols) block(ec, r, c + half, sub, blk);
            break;
        case PARTITION_SPLIT:
            partition(ec, r, c, sub, blk);
            partition(ec, r, c + half, sub, blk);
            partition(ec, r + half, c, sub, blk);
            partition(ec, r + half, c + half, sub, blk);
            break;
        case PARTITION_HORZ_A:
            block(ec, r, c, sp, blk); block(ec, r, c + half, sp, blk); block(ec, r + half, c, sub, blk);
            break;
        case PARTITION_HORZ_B:
            block(ec, r, c, sub, blk); block(ec, r + half, c, sp, blk); block(ec, r + half, c + half, sp, blk);
            break;
        case PARTITION_VERT_A:
            block(ec, r, c, sp, blk); block(ec, r + half, c, sp, blk); block(ec, r, c + half, sub, blk);
            break;
        case PARTITION_VERT_B:
            block(ec, r, c, sub, blk); block(ec, r, c + half, sp, blk); block(ec, r + half, c + half, sp, blk);
            break;
        case PARTITION_HORZ_4:
            for (int i = 0; i < 4; i++)
                if (i < 3 || r + quarter * 3 < MiRows) block(ec, r + quarter * i, c, sub, blk);
            break;
        default:   // PARTITION_VERT_4
            for (int i = 0; i < 4; i++)
                if (i < 3 || c + quarter * 3 < MiCols) block(ec, r, c + quarter * i, sub, blk);
            break;
        }
    }

    // decode_block() with intra_frame_mode_info() and residual() (5.11.5 - 5.11.39)
    void block(RangeEnc &ec, int r, int c, int bsize, size_t &blk) {
        if (!isAt(blk, r, c, bsize)) throw WriteError("the record's block order does not match its partition tree");
        const RecBlock &b = *(const RecBlock *)&R.blocks[blk];
        const uint32_t t0 = R.block_txb[blk], t1 = R.block_txb[blk + 1];
        blk++;
        const int availU = inside(r - 1, c), availL = inside(r, c - 1);
        const int hasC = has_chroma(bsize, r, c, mono, ss);
        // intra_frame_mode_info: skip, y mode, angle delta, uv mode (+ CfL), uv angle delta
        const int sctx = (availU ? at(mSkip, r - 1, c) : 0) + (availL ? at(mSkip, r, c - 1) : 0);
        ec.adapt(b.skip ? 1 : 0, cdf.skip[sctx], 2);
        // read_cdef: the superblock's index with its first coded block (cdef_bits literal bits), the same in the others
        if (cdefOn ? (b.cdef_idx < 0 || b.cdef_idx >= (1 << cdefBits)) : b.cdef_idx != 0) throw WriteError("cdef_idx out of range");
        if (cdefOn && !b.skip) {
            if (sbCdef < 0) {
                ec.literal((uint32_t)b.cdef_idx, cdefBits);
                sbCdef = b.cdef_idx;
            } else if (b.cdef_idx != sbCdef) {
                throw WriteError("cdef_idx differs from its superblock's");
            }
        }
        const int am = av1t::Intra_Mode_Context[availU ? at(mYMode, r - 1, c) : DC_PRED];
        const int lm = av1t::Intra_Mode_Context[availL ? at(mYMode, r, c - 1) : DC_PRED];
        if (b.y_mode > PAETH_PRED) throw WriteError("bad y_mode");
        ec.adapt(b.y_mode, cdf.yMode[am][lm], 13);
        if (bsize >= BLOCK_8X8 && is_dir(b.y_mode)) ec.adapt(b.angle_delta_y + 3, cdf.angle[b.y_mode - V_PRED], 7);
        if (hasC) {
            if (cfl_allowed(bsize)) ec.adapt(b.uv_mode, cdf.uvCfl[b.y_mode], 14);
            else {
                if (b.uv_mode == UV_CFL_PRED) throw WriteError("CfL in a block that does not allow it");
                ec.adapt(b.uv_mode, cdf.uvNoCfl[b.y_mode], 13);
            }
            if (b.uv_mode == UV_CFL_PRED) {
                const int sU = b.cfl_alpha_u == 0 ? 0 : b.cfl_alpha_u < 0 ? 1 : 2, sV = b.cfl_alpha_v == 0 ? 0 : b.cfl_alpha_v < 0 ? 1 : 2;
                if (sU == 0 && sV == 0) throw WriteError("CfL with both alphas zero");
                ec.adapt(sU * 3 + sV - 1, cdf.cflSign, 8);
                if (sU) ec.adapt(abs(b.cfl_alpha_u) - 1, cdf.cflAlpha[(sU - 1) * 3 + sV], 16);
                if (sV) ec.adapt(abs(b.cfl_alpha_v) - 1, cdf.cflAlpha[(sV - 1) * 3 + sU], 16);
            }
            if (bsize >= BLOCK_8X8 && is_dir(b.uv_mode))
                ec.adapt(b.angle_delta_uv + 3, cdf.angle[b.uv_mode - V_PRED], 7);
        }
        if (b.tx_depth != 0 || b.filter_intra != 0) throw WriteError("tx_depth / filter intra are not written");
        const int bh = bh4(bsize), bw = bw4(bsize);
        if (b.skip) {   // reset_block_context
            for (int p = 0; p < 1 + 2 * hasC; p++) {
                const int sp = p ? ss : 0;
                for (int i = c >> sp; i < ((c + bw) >> sp); i++) aLev[p][i] = aDc[p][i] = 0;
                for (int i = (r & 15) >> sp; i < (((r & 15) + bh) >> sp); i++) lLev[p][i] = lDc[p][i] = 0;
            }
        }
        const int rEnd = std::min(r + bh, MiRows), cEnd = std::min(c + bw, MiCols);
        for (int y = r; y < rEnd; y++)
            for (int x = c; x < cEnd; x++) at(mYMode, y, x) = b.y_mode;
        if (!b.skip) {
            for (uint32_t ti = t0; ti < t1; ti++) coeffs(ec, b, *(const RecTxb *)&R.txb[ti], bsize);
        } else {
            for (uint32_t ti = t0; ti < t1; ti++)
                if (R.txb[ti].eob) throw WriteError("a skip block with coefficients");
        }
        for (int y = r; y < rEnd; y++)
            for (int x = c; x < cEnd; x++) { at(mSkip, y, x) = b.skip; at(mSize, y, x) = (uint8_t)bsize; }
    }

    // coeffs() (5.11.39) of one tx block, and the level / dc-sign contexts it leaves
    void coeffs(RangeEnc &ec, const RecBlock &b, const RecTxb &t, int bsize) {
        const int plane = t.plane, txSz = t.tx_size, pt = plane > 0, sp = plane ? ss : 0;
        const int x4 = t.x4, y4 = t.y4, w4 = tx_w(txSz) >> 2, h4 = tx_h(txSz) >> 2;
        const int maxX4 = MiCols >> sp, maxY4 = MiRows >> sp;
        const int ctxSz = txsz_ctx_of(txSz);
        const int16_t *q = R.coeffs + t.coeff_off;
        std::vector<uint8_t> &aL = aLev[plane], &aD = aDc[plane], &lL = lLev[plane], &lD = lDc[plane];
        // the left arrays hold the current superblock row: plane 4x4 row modulo the superblock height
        auto lidx = [&](int k) { return (y4 + k) & ((16 >> sp) - 1); };
        int ctx;
        if (plane == 0) {
            int top = 0, left = 0;
            for (int k = 0; k < w4; k++) if (x4 + k < maxX4) top = std::max(top, (int)aL[x4 + k]);
            for (int k = 0; k < h4; k++) if (y4 + k < maxY4) left = std::max(left, (int)lL[lidx(k)]);
            const int mx = std::max(top, left), mn = std::min(top, left);
            if (4 * bw4(bsize) == tx_w(txSz) && 4 * bh4(bsize) == tx_h(txSz)) ctx = 0;
            else if (top == 0 && left == 0) ctx = 1;
            else if (top == 0 || left == 0) ctx = 2 + (mx > 3);
            else if (mx <= 3) ctx = 4;
            else if (mn <= 3) ctx = 5;
            else ctx = 6;
        } else {
            int above = 0, left = 0;
            for (int k = 0; k < w4; k++) if (x4 + k < maxX4) above |= aL[x4 + k] | aD[x4 + k];
            for (int k = 0; k < h4; k++) if (y4 + k < maxY4) left |= lL[lidx(k)] | lD[lidx(k)];
            const int ub = uv_bsize(bsize, ss);
            ctx = (above != 0) + (left != 0) + 7 + (16 * bw4(ub) * bh4(ub) > tx_w(txSz) * tx_h(txSz) ? 3 : 0);
        }
        const int eob = t.eob;
        ec.adapt(eob == 0, cdf.txbSkip[ctxSz][ctx], 2);
        int culLevel = 0, dcCategory = 0;
        if (eob) {
            int txType = t.tx_type;
            if (plane == 0) {
                const int set = tx_set_intra(txSz), sqr = av1t::Tx_Size_Sqr[txSz];
                if (set == TX_SET_INTRA_1) {
                    int idx = -1;
                    for (int i = 0; i < 7; i++) if (av1t::Tx_Type_Intra_Inv_Set1[i] == txType) idx = i;
                    if (idx < 0) throw WriteError("luma tx type outside its set");
                    ec.adapt(idx, cdf.set1[sqr][b.y_mode], 7);
                } else if (set == TX_SET_INTRA_2) {
                    int idx = -1;
                    for (int i = 0; i < 5; i++) if (av1t::Tx_Type_Intra_Inv_Set2[i] == txType) idx = i;
                    if (idx < 0) throw WriteError("luma tx type outside its set");
                    ec.adapt(idx, cdf.set2[sqr][b.y_mode], 5);
                } else if (txType != DCT_DCT) {
                    throw WriteError("a 32/64-point luma tx block must be DCT_DCT");
                }
            } else if (txType != chroma_tx_type(txSz, b.uv_mode)) {
                throw WriteError("chroma tx type differs from the one its mode implies");
            }
            const uint16_t *scan = scan_of(txSz, txType);
            const int cls = tx_class(txType), n = tx_ncoef(txSz);
            if (eob > n) throw WriteError("eob out of range");
            // eob_pt, eob_extra, eob bits
            const int lw = av1t::Tx_Width_Log2[txSz], lh = av1t::Tx_Height_Log2[txSz];
            const int ms = std::min(lw, 5) + std::min(lh, 5) - 4, ectx = cls == TX_CLASS_2D ? 0 : 1;
            const int eobPt = eob <= 2 ? eob : 2 + floor_log2(eob - 1);
            uint16_t *pc = ms == 0 ? cdf.eob16[pt][ectx] : ms == 1 ? cdf.eob32[pt][ectx] : ms == 2 ? cdf.eob64[pt][ectx]
                         : ms == 3 ? cdf.eob128[pt][ectx] : ms == 4 ? cdf.eob256[pt][ectx] : ms == 5 ? cdf.eob512[pt] : cdf.eob1024[pt];
            ec.adapt(eobPt - 1, pc, ms + 5);
            if (eobPt >= 3) {
                const int rem = eob - ((1 << (eobPt - 2)) + 1);
                ec.adapt((rem >> (eobPt - 3)) & 1, cdf.eobExtra[ctxSz][pt][eobPt - 3], 2);
                for (int i = eobPt - 4; i >= 0; i--) ec.bit((rem >> i) & 1);
            }
            // levels in reverse scan order with contexts from the levels already coded
            const int adj = av1t::Adjusted_Tx_Size[txSz], bwl = av1t::Tx_Width_Log2[adj];
            const int width = 1 << bwl, height = av1t::Tx_Height[adj];
            std::vector<int> lev((size_t)width * height, 0);
            for (int ci = eob - 1; ci >= 0; ci--) {
                const int pos = scan[ci], v = abs((int)q[pos]), lv = std::min(v, 15);
                const int row = pos >> bwl, col = pos - (row << bwl);
                if (ci == eob - 1) {
                    if (v == 0) throw WriteError("the level at eob - 1 is zero");
                    const int bctx = ci == 0 ? 0 : ci <= (height << bwl) / 8 ? 1 : ci <= (height << bwl) / 4 ? 2 : 3;
                    ec.adapt(std::min(lv, 3) - 1, cdf.baseEob[ctxSz][pt][bctx], 3);
                } else {
                    int mag = 0;
                    for (int k = 0; k < 5; k++) {
                        const int rr = row + av1t::Sig_Ref_Diff_Offset[cls][k][0], cc = col + av1t::Sig_Ref_Diff_Offset[cls][k][1];
                        if (rr < height && cc < width) mag += std::min(lev[(size_t)(rr << bwl) + cc], 3);
                    }
                    int bctx = std::min((mag + 1) >> 1, 4);
                    if (cls == TX_CLASS_2D) bctx = (row == 0 && col == 0) ? 0 : bctx + av1t::Coeff_Base_Ctx_Offset[txSz][std::min(row, 4)][std::min(col, 4)];
                    else bctx += av1t::Coeff_Base_Pos_Ctx_Offset[std::min(cls == TX_CLASS_VERT ? row : col, 2)];
                    ec.adapt(std::min(lv, 3), cdf.base[ctxSz][pt][bctx], 4);
                }
                if (lv > 2) {
                    int mag = 0;
                    for (int k = 0; k < 3; k++) {
                        const int rr = row + av1t::Mag_Ref_Offset_With_Tx_Class[cls][k][0], cc = col + av1t::Mag_Ref_Offset_With_Tx_Class[cls][k][1];
                        if (rr < height && cc < width) mag += std::min(lev[(size_t)rr * width + cc], 15);
                    }
                    mag = std::min((mag + 1) >> 1, 6);
                    const int brctx = pos == 0 ? mag : cls == TX_CLASS_2D ? ((row < 2 && col < 2) ? mag + 7 : mag + 14)
                                    : cls == TX_CLASS_HORIZ ? (col == 0 ? mag + 7 : mag + 14) : (row == 0 ? mag + 7 : mag + 14);
                    int rem = lv - 3;
                    for (int i = 0; i < 4; i++) {
                        const int k = std::min(rem, 3);
                        ec.adapt(k, cdf.br[std::min(ctxSz, (int)TX_32X32)][pt][brctx], 4);
                        rem -= k;
                        if (k < 3) break;
                    }
                }
                lev[pos] = lv;
            }
            // signs and Golomb remainders in scan order
            for (int ci = 0; ci < eob; ci++) {
                const int pos = scan[ci], qv = q[pos];
                if (!qv) continue;
                const int a = abs(qv);
                if (ci == 0) {
                    int s = 0;
                    for (int k = 0; k < w4; k++) if (x4 + k < maxX4) s += aD[x4 + k] == 2 ? 1 : aD[x4 + k] == 1 ? -1 : 0;
                    for (int k = 0; k < h4; k++) if (y4 + k < maxY4) s += lD[lidx(k)] == 2 ? 1 : lD[lidx(k)] == 1 ? -1 : 0;
                    ec.adapt(qv < 0, cdf.dcSign[pt][s < 0 ? 1 : s > 0 ? 2 : 0], 2);
                } else {
                    ec.bit(qv < 0);
                }
                if (a > 14) {   // read_golomb of a - 15
                    const uint32_t x = (uint32_t)(a - 15) + 1;
                    const int len = floor_log2((int)x) + 1;
                    for (int i = 0; i < len - 1; i++) ec.bit(0);
                    ec.bit(1);
                    for (int i = len - 2; i >= 0; i--) ec.bit((x >> i) & 1);
                }
                if (pos == 0) dcCategory = qv < 0 ? 1 : 2;
                culLevel += a;
            }
            culLevel = std::min(culLevel, 63);
        }
        for (int k = 0; k < w4; k++) { aL[x4 + k] = (uint8_t)culLevel; aD[x4 + k] = (uint8_t)dcCategory; }
        for (int k = 0; k < h4; k++) { lL[lidx(k)] = (uint8_t)culLevel; lD[lidx(k)] = (uint8_t)dcCategory; }
    }

    static int floor_log2(int v) {
        int k = 0;
        while ((v >> (k + 1)) > 0) k++;
        return k;
    }
};

// ---- AVIF (HEIF still image) container ----
struct Box {
    std::vector<uint8_t> d;
    void u8(uint32_t v) { d.push_back((uint8_t)v); }
    void u16(uint32_t v) { u8(v >> 8); u8(v); }
    void u32(uint32_t v) { u16(v >> 16); u16(v); }
    void str(const char *s, bool nul = false) { while (*s) u8((uint8_t)*s++); if (nul) u8(0); }
    void add(const std::vector<uint8_t> &v) { d.insert(d.end(), v.begin(), v.end()); }
};
std::vector<uint8_t> box(const char *type, const std::vector<uint8_t> &payload, int version = -1, uint32_t flags = 0) {
    Box b;
    b.u32((uint32_t)(8 + (version >= 0 ? 4 : 0) + payload.size()));
    b.str(type);
    if (version >= 0) { b.u8((uint32_t)version); b.u8(flags >> 16); b.u16(flags & 0xFFFF); }
    b.add(payload);
    return b.d;
}

std::vector<uint8_t> avif(const std::vector<uint8_t> &obus, const std::vector<uint8_t> &seqObu, int W, int H, int bd, int mono,
                          int level, int profile, int ss) {
    Box ftyp;
    ftyp.str("avif"); ftyp.u32(0); ftyp.str("avif"); ftyp.str("mif1"); ftyp.str("miaf");
    Box hdlr;
    hdlr.u32(0); hdlr.str("pict"); hdlr.u32(0); hdlr.u32(0); hdlr.u32(0); hdlr.str("", true);
    Box pitm;
    pitm.u16(1);
    Box infe;
    infe.u16(1); infe.u16(0); infe.str("av01"); infe.str("", true);
    Box iinf;
    iinf.u16(1); iinf.add(box("infe", infe.d, 2));
    Box ispe;
    ispe.u32((uint32_t)W); ispe.u32((uint32_t)H);
    Box av1c;
    av1c.u8(0x81);                                   // marker, version 1
    av1c.u8(((uint32_t)profile << 5) | (uint32_t)level);   // seq_profile, seq_level_idx_0
    av1c.u8(((bd == 10) << 6) | (mono << 4) | (ss << 3) | (ss << 2));   // tier 0, high_bitdepth, monochrome, subsampling x / y
    av1c.u8(0);                                      // no initial presentation delay
    av1c.add(seqObu);
    Box pixi;
    const int nch = mono ? 1 : 3;
    pixi.u8((uint32_t)nch);
    for (int i = 0; i < nch; i++) pixi.u8((uint32_t)bd);
    Box colr;
    colr.str("nclx"); colr.u16(1); colr.u16(13); colr.u16(6); colr.u8(0x80);   // BT.709 / sRGB / BT.601, full range
    Box ipco;
    ipco.add(box("ispe", ispe.d, 0)); ipco.add(box("av1C", av1c.d)); ipco.add(box("pixi", pixi.d, 0)); ipco.add(box("colr", colr.d));
    Box ipma;
    ipma.u32(1); ipma.u16(1); ipma.u8(4);
    ipma.u8(1); ipma.u8(0x80 | 2); ipma.u8(3); ipma.u8(0x80 | 4);   // ispe, av1C (essential), pixi, colr (essential)
    Box iprp;
    iprp.add(box("ipco", ipco.d)); iprp.add(box("ipma", ipma.d, 0));
    // iloc points into mdat: its offset is known once the meta box's size is
    auto meta_with = [&](uint32_t offset) {
        Box iloc;
        iloc.u8(0x44); iloc.u8(0x00);   // offset_size 4, length_size 4, base_offset_size 0, reserved
        iloc.u16(1); iloc.u16(1); iloc.u16(0); iloc.u16(1); iloc.u32(offset); iloc.u32((uint32_t)obus.size());
        Box meta;
        meta.add(box("hdlr", hdlr.d, 0)); meta.add(box("pitm", pitm.d, 0)); meta.add(box("iloc", iloc.d, 0));
        meta.add(box("iinf", iinf.d, 0)); meta.add(box("iprp", iprp.d));
        return box("meta", meta.d, 0);
    };
    const std::vector<uint8_t> f = box("ftyp", ftyp.d);
    const size_t metaLen = meta_with(0).size();
    const uint32_t offset = (uint32_t)(f.size() + metaLen + 8);
    std::vector<uint8_t> out = f;
    const std::vector<uint8_t> meta = meta_with(offset);
    out.insert(out.end(), meta.begin(), meta.end());
    const std::vector<uint8_t> mdat = box("mdat", obus);
    out.insert(out.end(), mdat.begin(), mdat.end());
    return out;
}

}  // namespace

extern "C" {

int av1mi_write_obus(const av1mi_frame_out *rec, int edgeFilter, void **data, size_t *size, char *err, int errlen) {
    try {
        Writer w(*rec, edgeFilter);
        const std::vector<uint8_t> o = w.obus();
        *data = malloc(o.size());
        memcpy(*data, o.data(), o.size());
        *size = o.size();
        return 0;
    } catch (const std::exception &e) {
        snprintf(err, errlen, "%s", e.what());
        return -1;
    }
}

int av1mi_write_avif(const av1mi_frame_out *rec, int edgeFilter, void **data, size_t *size, char *err, int errlen) {
    try {
        Writer w(*rec, edgeFilter);
        const std::vector<uint8_t> o = w.obus();
        const std::vector<uint8_t> f = avif(o, w.seqObu, (int)rec->width, (int)rec->height, rec->bit_depth == 10 ? 10 : 8,
                                            rec->monochrome ? 1 : 0, w.level(), w.profile(), w.subsampling());
        *data = malloc(f.size());
        memcpy(*data, f.data(), f.size());
        *size = f.size();
        return 0;
    } catch (const std::exception &e) {
        snprintf(err, errlen, "%s", e.what());
        return -1;
    }
}

void av1mi_buffer_free(void *p) { free(p); }

}  // extern "C"
