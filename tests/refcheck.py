"""Test infrastructure: ctypes bindings of the checkers.

* oracle/liborc.so      -- the CPU restatement of the reference path (the parity oracle).
* oracle/_ref/libav1ref.so -- the reference's own CPU backend (av1core encoder.cpp + avif.cpp) and header-only DSP,
  compiled from the reference sources by oracle/Makefile: decision record -> decoder-exact reconstruction, OBUs, AVIF.
* Pillow's bundled libavif (dav1d decoder) for the AVIF decode check.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
"""
import ctypes as C
import glob
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_SO = os.path.join(ROOT, "oracle", "liborc.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libav1ref.so")

BLOCK_DTYPE = np.dtype([("mi_row", "<u2"), ("mi_col", "<u2"), ("bsize", "u1"), ("skip", "u1"), ("tx_depth", "u1"),
                        ("y_mode", "u1"), ("angle_delta_y", "i1"), ("uv_mode", "u1"), ("angle_delta_uv", "i1"),
                        ("cfl_alpha_u", "i1"), ("cfl_alpha_v", "i1"), ("filter_intra", "u1"), ("cdef_idx", "i1"),
                        ("sb_qindex", "u1")])
TXB_DTYPE = np.dtype([("coeff_off", "<u4"), ("x4", "<u2"), ("y4", "<u2"), ("plane", "u1"), ("tx_size", "u1"),
                      ("tx_type", "u1"), ("flags", "u1"), ("eob", "<u2"), ("reserved", "<u2")])
BLOCKRES_DTYPE = np.dtype([("J", "<f4"), ("cand", "u1", 6), ("ncand", "u1"), ("tx", "u1")])
LEAF_DTYPE = np.dtype([("bsize", "u1"), ("x4", "u1"), ("y4", "u1"), ("pad", "u1")])
assert BLOCK_DTYPE.itemsize == 16 and TXB_DTYPE.itemsize == 16 and BLOCKRES_DTYPE.itemsize == 12


EXT_DTYPE = np.dtype([("J", "<f4"), ("mode", "u1"), ("delta", "i1"), ("tx", "u1"), ("pad", "u1")])   # GpExtCand


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("lambdaMul", "satdLambda", "rnd", "zbin", "rndC", "zbinC")] + \
               [(n, C.c_int) for n in ("K", "Krecheck", "txSearch", "txSearchCL", "txAllCL", "deltas", "uvDeltas", "cfl",
                                      "uvK", "shapeMask", "partMask8", "partMaskN", "maxPart")] + \
               [(n, C.c_float) for n in ("uvOverheadBits", "splitBias", "lambdaCMul")]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self):
        self.lib = C.CDLL(ORC_SO)
        L = self.lib
        L.orc_encode.restype = C.c_void_p
        L.orc_encode.argtypes = [C.c_void_p] * 3 + [C.POINTER(C.c_int)] + [C.c_int] * 7 + [C.POINTER(OrcParams), C.c_int]
        L.orc_encode_hooks.restype = C.c_void_p
        L.orc_encode_hooks.argtypes = [C.c_void_p] * 3 + [C.POINTER(C.c_int)] + [C.c_int] * 7 + [C.POINTER(OrcParams), C.c_int,
                                                                                          C.c_void_p, C.c_int, C.c_void_p]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_counts.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_record.argtypes = [C.c_void_p] * 5
        L.orc_recon.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_blockres.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_leaves.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_sbj.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_fwd_matrix.argtypes = [C.c_int, C.c_int, C.c_void_p]
        L.orc_misc_tables.argtypes = [C.c_void_p] * 7
        L.orc_cost_table.argtypes = [C.c_int, C.c_void_p]
        L.orc_eval_txb.argtypes = [C.c_void_p] + [C.c_int] * 8 + [C.c_float, C.c_float] + [C.c_void_p] * 3
        L.orc_rate_levels.restype = C.c_float
        L.orc_rate_levels.argtypes = [C.c_void_p] + [C.c_int] * 7 + [C.POINTER(C.c_int)]
        L.orc_dequant_inverse.restype = C.c_uint32
        L.orc_dequant_inverse.argtypes = [C.c_void_p] + [C.c_int] * 5 + [C.c_void_p]
        L.orc_predict.argtypes = [C.c_int] * 11 + [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_build_edges.argtypes = [C.c_void_p] + [C.c_int] * 11 + [C.c_void_p, C.c_void_p]
        L.orc_tile_layout.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 4

    def rate_levels(self, q, tx_size, tx_type, plane, qindex, skip_ctx, dc_ctx, intra_dir):
        """(R in bits, eob) of the levels q (int16, row-major) of one tx block under the oracle's rate model."""
        q = np.ascontiguousarray(q, dtype=np.int16)
        eob = C.c_int()
        r = self.lib.orc_rate_levels(_ptr(q), tx_size, tx_type, plane, qindex, skip_ctx, dc_ctx, intra_dir, C.byref(eob))
        return float(r), eob.value

    def params(self, **kw):
        p = OrcParams()
        self.lib.orc_default_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def encode(self, planes, W, H, ssx=1, mono=0, bd=8, qindex=120, edge_filter=1, params=None, threads=1, detail=False,
               ext=None, ext_topk=0, analyse=None):
        """ext: {shape: EXT_DTYPE array [blocks of the frame's grid (raster), ext_topk]} external candidates;
        analyse: {shape: uint8 array [blocks of the grid]} 1 = analysed (the analysis hooks)."""
        strides = (C.c_int * 3)(*[p.shape[1] for p in planes])
        pl = [np.ascontiguousarray(p) for p in planes]
        keep = []
        ext_p = an_p = None
        if ext:
            ext_p = (C.c_void_p * 22)()
            for sh, a in ext.items():
                a = np.ascontiguousarray(a, EXT_DTYPE)
                keep.append(a)
                ext_p[sh] = a.ctypes.data
        if analyse:
            an_p = (C.c_void_p * 22)()
            for sh, a in analyse.items():
                a = np.ascontiguousarray(a, np.uint8)
                keep.append(a)
                an_p[sh] = a.ctypes.data
        h = self.lib.orc_encode_hooks(_ptr(pl[0]), _ptr(pl[1]), _ptr(pl[2]), strides, W, H, ssx, mono, bd, qindex, edge_filter,
                                      C.byref(params or self.params()), threads, ext_p, ext_topk, an_p)
        try:
            cnt = np.zeros(15, np.int64)
            self.lib.orc_counts(h, _ptr(cnt))
            nb, nt, nc, mic, mir, aw, ah = (int(x) for x in cnt[:7])
            out = dict(blocks=np.zeros(nb, BLOCK_DTYPE), block_txb=np.zeros(nb + 1, np.uint32), txb=np.zeros(nt, TXB_DTYPE),
                       coeffs=np.zeros(nc, np.int16), MiCols=mic, MiRows=mir, aw=aw, ah=ah, stats=cnt[9:13].copy(),
                       flops=(int(cnt[13]), int(cnt[14])))
            self.lib.orc_record(h, _ptr(out["blocks"]), _ptr(out["block_txb"]), _ptr(out["txb"]), _ptr(out["coeffs"]))
            rec = []
            for p in range(3):
                sx = ssx if p else 0
                r = np.zeros(((ah >> sx), (aw >> sx)), np.uint16)
                self.lib.orc_recon(h, p, _ptr(r))
                rec.append(r)
            out["recon"] = rec
            if detail:
                br = {}
                for s in list(range(13)) + list(range(16, 22)):
                    n = 64 * 64 * 16 if False else (((mic * 4 + 63) // 4) * ((mir * 4 + 63) // 4))
                    a = np.zeros(n, BLOCKRES_DTYPE)
                    k = self.lib.orc_blockres(h, s, _ptr(a))
                    br[s] = a[:k]
                out["blockres"] = br
                nsb = ((mic + 15) >> 4) * ((mir + 15) >> 4)
                cnts = np.zeros(nsb, np.uint16)
                lv = np.zeros(nsb * 256, LEAF_DTYPE)
                self.lib.orc_leaves(h, _ptr(cnts), _ptr(lv))
                out["leaf_count"] = cnts
                out["leaves"] = lv.reshape(nsb, 256)
                sbj = np.zeros(nsb, np.float32)
                self.lib.orc_sbj(h, _ptr(sbj))
                out["sbJ"] = sbj
            return out
        finally:
            self.lib.orc_free(h)


# ---------------------------------------------------------------------------------------------------------------
# the reference's CPU backend (oracle/_ref)
class FrameParams(C.Structure):
    _fields_ = [("version", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("monochrome", C.c_uint8),
                ("base_q_idx", C.c_uint8), ("tx_mode_select", C.c_uint8), ("reduced_tx_set", C.c_uint8),
                ("tile_cols_log2", C.c_int8), ("tile_rows_log2", C.c_int8), ("delta_q_y_dc", C.c_int8),
                ("delta_q_u_dc", C.c_int8), ("delta_q_u_ac", C.c_int8), ("delta_q_v_dc", C.c_int8), ("delta_q_v_ac", C.c_int8),
                ("enable_filter_intra", C.c_uint8), ("enable_intra_edge_filter", C.c_uint8), ("delta_q_present", C.c_uint8),
                ("delta_q_res", C.c_uint8), ("using_qmatrix", C.c_uint8), ("qm_y", C.c_uint8), ("qm_u", C.c_uint8),
                ("qm_v", C.c_uint8), ("loop_filter_level", C.c_uint8 * 4), ("loop_filter_sharpness", C.c_uint8),
                ("enable_cdef", C.c_uint8), ("cdef_damping_minus_3", C.c_uint8), ("cdef_bits", C.c_uint8),
                ("cdef_y_pri", C.c_uint8 * 8), ("cdef_y_sec", C.c_uint8 * 8), ("cdef_uv_pri", C.c_uint8 * 8),
                ("cdef_uv_sec", C.c_uint8 * 8), ("allow_screen_content_tools", C.c_uint8), ("chroma_format", C.c_uint8),
                ("bit_depth", C.c_uint8), ("color_description", C.c_uint8), ("color_primaries", C.c_uint8),
                ("transfer_characteristics", C.c_uint8), ("matrix_coefficients", C.c_uint8), ("reserved", C.c_uint8 * 9)]


class DecisionRecord(C.Structure):
    _fields_ = [("fp", FrameParams), ("n_blocks", C.c_uint32), ("n_txb", C.c_uint32), ("n_coeffs", C.c_uint64),
                ("blocks", C.c_void_p), ("block_txb", C.c_void_p), ("txb", C.c_void_p), ("coeffs", C.c_void_p),
                ("n_palettes", C.c_uint32), ("palettes", C.c_void_p), ("color_map", C.c_void_p), ("n_color_map", C.c_uint64)]


class EncodeOpts(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("seq_level_idx", C.c_int32)]


class Output(C.Structure):
    _fields_ = [("obu", C.c_void_p), ("obu_size", C.c_size_t), ("avif", C.c_void_p), ("avif_size", C.c_size_t),
                ("recon", C.c_void_p * 3), ("recon_w", C.c_uint32 * 3), ("recon_h", C.c_uint32 * 3),
                ("seq_level_idx", C.c_uint32), ("tile_cols", C.c_uint32), ("tile_rows", C.c_uint32),
                ("header_bytes", C.c_uint32), ("tile_bytes", C.c_uint64), ("n_symbols", C.c_uint64),
                ("n_fixed_txb", C.c_uint32), ("error", C.c_char * 256)]


class Tiles(C.Structure):
    _fields_ = [("cols_log2", C.c_int32), ("rows_log2", C.c_int32), ("cols", C.c_uint32), ("rows", C.c_uint32),
                ("mi_col_starts", C.c_uint32 * 65), ("mi_row_starts", C.c_uint32 * 65)]


AV1G_ENC_AVIF = 8
AV1G_ENC_PREFILTER_RECON = 0x20


def ref_available():
    return os.path.exists(REF_SO)


class Parsed(C.Structure):
    _fields_ = [("rec", DecisionRecord), ("seq_level_idx", C.c_uint32), ("n_tiles", C.c_uint32), ("tile_bytes", C.c_void_p),
                ("tile_offsets", C.c_void_p), ("error", C.c_char * 256), ("priv", C.c_void_p)]


class Ref:
    def __init__(self):
        self.lib = C.CDLL(REF_SO)
        L = self.lib
        L.av1g_encode.argtypes = [C.POINTER(DecisionRecord), C.POINTER(EncodeOpts), C.POINTER(Output)]
        L.av1g_output_free.argtypes = [C.POINTER(Output)]
        L.av1g_layout.argtypes = [C.POINTER(DecisionRecord), C.c_uint32, C.c_uint64]
        L.av1g_validate.argtypes = [C.POINTER(DecisionRecord), C.c_char_p, C.c_size_t]
        L.av1g_tile_layout.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.POINTER(Tiles)]
        L.ref_dequant_inverse.restype = C.c_uint32
        L.ref_dequant_inverse.argtypes = [C.c_void_p] + [C.c_int] * 5 + [C.c_void_p]
        L.ref_predict.argtypes = [C.c_int] * 11 + [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.ref_build_edges.argtypes = [C.c_void_p] + [C.c_int] * 11 + [C.c_void_p, C.c_void_p]
        L.av1g_parse.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.POINTER(Parsed))]
        L.av1g_parsed_free.argtypes = [C.POINTER(Parsed)]
        L.ref_tile_symbols.restype = C.c_long
        L.ref_tile_symbols.argtypes = [C.POINTER(DecisionRecord), C.c_void_p, C.c_size_t, C.c_char_p, C.c_size_t]
        L.ref_default_cdfs.restype = C.c_size_t
        L.ref_default_cdfs.argtypes = [C.c_int, C.c_void_p]
        L.ref_coeff_cdf_fields.argtypes = [C.c_void_p]

    COEFF_CDF_FIELDS = ("intra_tx_set1", "intra_tx_set2", "txb_skip", "eob_pt_16", "eob_pt_32", "eob_pt_64", "eob_pt_128",
                        "eob_pt_256", "eob_pt_512", "eob_pt_1024", "eob_extra", "dc_sign", "coeff_base_eob", "coeff_base",
                        "coeff_br")

    def coeff_cdf_fields(self):
        """{field: (cdf id of its first entry, size in uint16)} of the CDFs the coefficient syntax uses (cdfctx.h)."""
        a = np.zeros(2 * len(self.COEFF_CDF_FIELDS), np.uint32)
        assert self.lib.ref_coeff_cdf_fields(_ptr(a)) == len(self.COEFF_CDF_FIELDS)
        return {f: (int(a[2 * i]), int(a[2 * i + 1])) for i, f in enumerate(self.COEFF_CDF_FIELDS)}

    def default_cdfs(self, qindex):
        """The uint16 CDF state a tile of a frame with base_q_idx = qindex starts from (av1g_cdfs_init)."""
        n = self.lib.ref_default_cdfs(qindex, None)
        a = np.zeros(n, np.uint16)
        self.lib.ref_default_cdfs(qindex, _ptr(a))
        return a

    def layout(self, blocks, W, H, qindex, ssx=1, mono=0, bd=8):
        """av1g_layout of a block list: (block_txb, txb) with coeff_off / x4 / y4 / plane / tx_size filled."""
        blocks = np.ascontiguousarray(blocks)
        r = DecisionRecord()
        r.fp = self.frame_params(W, H, qindex, ssx, mono, bd)
        r.n_blocks = len(blocks)
        r.blocks = _ptr(blocks)
        assert self.lib.av1g_layout(C.byref(r), 0, 0) != 0   # counts only
        bt = np.zeros(len(blocks) + 1, np.uint32)
        txb = np.zeros(r.n_txb, TXB_DTYPE)
        r.block_txb, r.txb = _ptr(bt), _ptr(txb)
        rc = self.lib.av1g_layout(C.byref(r), len(txb), r.n_coeffs)
        if rc != 0:
            raise AssertionError("av1g_layout rc %d" % rc)
        return bt, txb, int(r.n_coeffs)

    def symbols(self, rec, W, H, qindex, ssx=1, mono=0, bd=8, edge_filter=1):
        """The reference writer's symbol stream of a record (all tiles, in order) as av1g_sym_pack words."""
        arrs = [np.ascontiguousarray(rec[k]).copy() for k in ("blocks", "block_txb", "txb", "coeffs")]
        r = DecisionRecord()
        r.fp = self.frame_params(W, H, qindex, ssx, mono, bd, edge_filter)
        r.n_blocks, r.n_txb, r.n_coeffs = len(arrs[0]), len(arrs[2]), len(arrs[3])
        r.blocks, r.block_txb, r.txb, r.coeffs = (_ptr(a) for a in arrs)
        err = C.create_string_buffer(256)
        n = self.lib.ref_tile_symbols(C.byref(r), None, 0, err, 256)
        if n < 0:
            raise AssertionError("ref_tile_symbols: " + err.value.decode())
        out = np.zeros(n, np.uint32)
        self.lib.ref_tile_symbols(C.byref(r), _ptr(out), n, err, 256)
        return out

    def parse(self, obus):
        """The reference's OBU parser (av1core/src/encoder.cpp:1683-1969): the record a stream codes, as a dict."""
        p = C.POINTER(Parsed)()
        if self.lib.av1g_parse(obus, len(obus), C.byref(p)) != 0:
            msg = p.contents.error.decode() if p else "?"
            if p:
                self.lib.av1g_parsed_free(p)
            raise AssertionError("av1g_parse: " + msg)
        try:
            r = p.contents.rec
            nb, nt, nc = r.n_blocks, r.n_txb, r.n_coeffs
            out = dict(blocks=np.frombuffer(C.string_at(r.blocks, nb * 16), BLOCK_DTYPE).copy(),
                       block_txb=np.frombuffer(C.string_at(r.block_txb, (nb + 1) * 4), np.uint32).copy(),
                       txb=np.frombuffer(C.string_at(r.txb, nt * 16), TXB_DTYPE).copy(),
                       coeffs=np.frombuffer(C.string_at(r.coeffs, nc * 2), np.int16).copy(),
                       base_q_idx=r.fp.base_q_idx, width=r.fp.width, height=r.fp.height, bit_depth=r.fp.bit_depth,
                       edge_filter=r.fp.enable_intra_edge_filter, n_tiles=p.contents.n_tiles,
                       post={f: (list(getattr(r.fp, f)) if hasattr(getattr(r.fp, f), "__len__") else getattr(r.fp, f))
                             for f in self.PF_FIELDS})
            return out
        finally:
            self.lib.av1g_parsed_free(p)

    @staticmethod
    def frame_params(W, H, qindex, ssx=1, mono=0, bd=8, edge_filter=1):
        fp = FrameParams()
        fp.version = 5
        fp.width, fp.height = W, H
        fp.monochrome = mono
        fp.base_q_idx = qindex
        fp.tile_cols_log2 = -1
        fp.tile_rows_log2 = -1
        fp.enable_intra_edge_filter = edge_filter
        fp.chroma_format = 0 if (ssx or mono) else 1
        fp.bit_depth = bd
        return fp

    PF_FIELDS = ("loop_filter_level", "loop_filter_sharpness", "enable_cdef", "cdef_damping_minus_3", "cdef_bits", "cdef_y_pri",
                 "cdef_y_sec", "cdef_uv_pri", "cdef_uv_sec")

    def encode(self, rec, W, H, qindex, ssx=1, mono=0, bd=8, edge_filter=1, avif=False, layout_check=True, post=None):
        """Decoder-exact reconstruction (and AVIF bytes) of a record; returns (recon planes cropped, avif bytes, info).
        post: frame parameters whose loop filter / CDEF fields the frame takes (the reconstruction is then the
        post-filtered one, as the reference's decoder-exact backend applies them)."""
        blocks = np.ascontiguousarray(rec["blocks"])
        block_txb = np.ascontiguousarray(rec["block_txb"]).astype(np.uint32)
        txb = np.ascontiguousarray(rec["txb"]).copy()
        coeffs = np.ascontiguousarray(rec["coeffs"]).copy()
        r = DecisionRecord()
        r.fp = self.frame_params(W, H, qindex, ssx, mono, bd, edge_filter)
        if post is not None:
            for f in self.PF_FIELDS:
                v = getattr(post, f)
                if hasattr(v, "__len__"):
                    for k in range(len(v)):
                        getattr(r.fp, f)[k] = v[k]
                else:
                    setattr(r.fp, f, v)
        r.n_blocks = len(blocks)
        if layout_check:   # av1g_layout must reproduce the engine's tx-block layout exactly
            bt2 = np.zeros(len(blocks) + 1, np.uint32)
            t2 = np.zeros(len(txb), TXB_DTYPE)
            r.blocks, r.block_txb, r.txb = _ptr(blocks), _ptr(bt2), _ptr(t2)
            rc = self.lib.av1g_layout(C.byref(r), len(t2), len(coeffs))
            if rc != 0:
                raise AssertionError("av1g_layout rc %d (n_txb %d vs %d)" % (rc, r.n_txb, len(txb)))
            if not np.array_equal(bt2, block_txb):
                raise AssertionError("block_txb differs from av1g_layout")
            for f in ("coeff_off", "x4", "y4", "plane", "tx_size"):
                if not np.array_equal(t2[f], txb[f]):
                    raise AssertionError("txb.%s differs from av1g_layout" % f)
        r.blocks, r.block_txb, r.txb, r.coeffs = _ptr(blocks), _ptr(block_txb), _ptr(txb), _ptr(coeffs)
        r.n_txb = len(txb)
        r.n_coeffs = len(coeffs)
        err = C.create_string_buffer(256)
        if self.lib.av1g_validate(C.byref(r), err, 256) != 0:
            raise AssertionError("av1g_validate: " + err.value.decode())
        opts = EncodeOpts(AV1G_ENC_AVIF if avif else 0, -1)
        out = Output()
        rc = self.lib.av1g_encode(C.byref(r), C.byref(opts), C.byref(out))
        if rc != 0:
            raise AssertionError("av1g_encode rc %d: %s" % (rc, out.error.decode()))
        try:
            planes = []
            bps = 2 if bd > 8 else 1
            for p in range(1 if mono else 3):
                w, h = out.recon_w[p], out.recon_h[p]
                buf = C.string_at(out.recon[p], w * h * bps)
                planes.append(np.frombuffer(buf, np.uint16 if bps == 2 else np.uint8).reshape(h, w).copy())
            av = C.string_at(out.avif, out.avif_size) if avif else None
            info = dict(obu_size=out.obu_size, n_fixed=out.n_fixed_txb, eob=txb["eob"].copy(), tx_type=txb["tx_type"].copy())
            return planes, av, info
        finally:
            self.lib.av1g_output_free(C.byref(out))

    def tile_layout(self, W, H):
        t = Tiles()
        assert self.lib.av1g_tile_layout(W, H, -1, -1, C.byref(t)) == 0
        return list(t.mi_col_starts[:t.cols + 1]), list(t.mi_row_starts[:t.rows + 1])


# ---------------------------------------------------------------------------------------------------------------
# dav1d through Pillow's bundled libavif: decode an AVIF into its YUV planes
class _AvifImage(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("depth", C.c_uint32), ("yuvFormat", C.c_int),
                ("yuvRange", C.c_int), ("yuvChromaSamplePosition", C.c_int), ("yuvPlanes", C.c_void_p * 3),
                ("yuvRowBytes", C.c_uint32 * 3)]


def _libavif():
    import PIL
    libs = glob.glob(os.path.join(os.path.dirname(PIL.__file__), "..", "pillow.libs", "libavif-*.so*"))
    return C.CDLL(libs[0]) if libs else None


def dav1d_available():
    try:
        from PIL import _avif
        return _avif.decoder_codec_available("dav1d") and _libavif() is not None
    except Exception:
        return False


def dav1d_decode_yuv(avif_bytes):
    """YUV planes (numpy, uint8 or uint16) of an AVIF decoded by libavif's dav1d."""
    L = _libavif()
    L.avifDecoderCreate.restype = C.c_void_p
    L.avifImageCreateEmpty.restype = C.POINTER(_AvifImage)
    L.avifDecoderReadMemory.argtypes = [C.c_void_p, C.POINTER(_AvifImage), C.c_char_p, C.c_size_t]
    L.avifDecoderDestroy.argtypes = [C.c_void_p]
    L.avifImageDestroy.argtypes = [C.POINTER(_AvifImage)]
    dec = L.avifDecoderCreate()
    img = L.avifImageCreateEmpty()
    try:
        rc = L.avifDecoderReadMemory(dec, img, avif_bytes, len(avif_bytes))
        if rc != 0:
            raise AssertionError("avifDecoderReadMemory: %d" % rc)
        im = img.contents
        bps = 2 if im.depth > 8 else 1
        fmt = im.yuvFormat   # 1 = 444, 3 = 420, 4 = 400
        planes = []
        for p in range(1 if fmt == 4 else 3):
            if p == 0:
                w, h = im.width, im.height
            elif fmt == 3:
                w, h = (im.width + 1) // 2, (im.height + 1) // 2
            else:
                w, h = im.width, im.height
            rb = im.yuvRowBytes[p]
            buf = C.string_at(im.yuvPlanes[p], rb * h)
            a = np.frombuffer(buf, np.uint8).reshape(h, rb)[:, :w * bps]
            planes.append(a.view(np.uint16).copy() if bps == 2 else a.copy())
        return planes
    finally:
        L.avifImageDestroy(img)
        L.avifDecoderDestroy(dec)
