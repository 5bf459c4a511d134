"""av1gpu_amd -- MI355X-native AV1 intra rate-distortion core.

Python host mirror of the reference's encoder API (maddog107/av1gpu encoder/src/pipeline.h:191-216): ``GpEncoder``
with ``init(config)`` / ``encode(frames)`` / ``config()``, ``GpConfig`` / ``GpParams`` knobs, ``GpFrameIn`` /
``GpFrameOut`` / ``GpTimes``, over the C ABI of ``include/av1mi.h`` (``libav1mi.so``, built in-tree by
``av1gpu_amd/csrc/Makefile``).  The HIP library is the only compute path: if it is missing, every call raises.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libav1mi.so")

BLOCK_DTYPE = np.dtype([("mi_row", "<u2"), ("mi_col", "<u2"), ("bsize", "u1"), ("skip", "u1"), ("tx_depth", "u1"),
                        ("y_mode", "u1"), ("angle_delta_y", "i1"), ("uv_mode", "u1"), ("angle_delta_uv", "i1"),
                        ("cfl_alpha_u", "i1"), ("cfl_alpha_v", "i1"), ("filter_intra", "u1"), ("cdef_idx", "i1"),
                        ("sb_qindex", "u1")])
TXB_DTYPE = np.dtype([("coeff_off", "<u4"), ("x4", "<u2"), ("y4", "<u2"), ("plane", "u1"), ("tx_size", "u1"),
                      ("tx_type", "u1"), ("flags", "u1"), ("eob", "<u2"), ("reserved", "<u2")])

BLOCKRES_DTYPE = np.dtype([("J", "<f4"), ("cand", "u1", 6), ("ncand", "u1"), ("tx", "u1")])
LEAF_DTYPE = np.dtype([("bsize", "u1"), ("x4", "u1"), ("y4", "u1"), ("pad", "u1")])
SHAPES = [s for s in range(22) if not 13 <= s <= 15]   # the 19 block sizes the analysis evaluates

_PARAM_FLOATS = ("lambdaMul", "satdLambda", "rnd", "zbin", "rndC", "zbinC")
_PARAM_INTS = ("K", "Krecheck", "txSearch", "txSearchCL", "txAllCL", "deltas", "uvDeltas", "cfl", "uvK", "shapeMask",
               "partMask8", "partMaskN", "maxPart")
_PARAM_FLOATS2 = ("uvOverheadBits", "splitBias", "lambdaCMul")
_POSTFILTER_INTS = ("cdef", "cdefDampingMinus3", "cdefYPri", "cdefYSec", "cdefUVPri", "cdefUVSec")


class _Params(C.Structure):
    _fields_ = [(n, C.c_float) for n in _PARAM_FLOATS] + [(n, C.c_int) for n in _PARAM_INTS] + \
               [(n, C.c_float) for n in _PARAM_FLOATS2]


class _Config(C.Structure):
    _fields_ = [("qindex", C.c_int), ("edgeFilter", C.c_int), ("P", _Params), ("schedule", C.c_int), ("queueBlocks", C.c_int),
                ("anStreams", C.c_int), ("graph", C.c_int), ("tune", C.c_int), ("lfMode", C.c_int), ("lfMul", C.c_float),
                ("lfLevel", C.c_int)] + [(n, C.c_int) for n in _POSTFILTER_INTS] + [("queueOcc", C.c_int)]


class _FrameParams(C.Structure):
    """av1mi_frame_params = av1g_frame_params (decisions.h:269-313), 88 bytes."""
    _fields_ = [("version", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32)] + \
               [(n, C.c_uint8) for n in ("monochrome", "base_q_idx", "tx_mode_select", "reduced_tx_set")] + \
               [(n, C.c_int8) for n in ("tile_cols_log2", "tile_rows_log2", "delta_q_y_dc", "delta_q_u_dc", "delta_q_u_ac",
                                        "delta_q_v_dc", "delta_q_v_ac")] + \
               [(n, C.c_uint8) for n in ("enable_filter_intra", "enable_intra_edge_filter", "delta_q_present", "delta_q_res",
                                         "using_qmatrix", "qm_y", "qm_u", "qm_v")] + \
               [("loop_filter_level", C.c_uint8 * 4), ("loop_filter_sharpness", C.c_uint8), ("enable_cdef", C.c_uint8),
                ("cdef_damping_minus_3", C.c_uint8), ("cdef_bits", C.c_uint8), ("cdef_y_pri", C.c_uint8 * 8),
                ("cdef_y_sec", C.c_uint8 * 8), ("cdef_uv_pri", C.c_uint8 * 8), ("cdef_uv_sec", C.c_uint8 * 8)] + \
               [(n, C.c_uint8) for n in ("allow_screen_content_tools", "chroma_format", "bit_depth", "color_description",
                                         "color_primaries", "transfer_characteristics", "matrix_coefficients")] + \
               [("reserved", C.c_uint8 * 9)]


class _FrameIn(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("mono", C.c_int), ("yuv444", C.c_int), ("bitDepth", C.c_int),
                ("dsrc", C.c_void_p * 3), ("spitch", C.c_int * 3), ("aw", C.c_int), ("ah", C.c_int), ("qindex", C.c_int)]


class _FrameOut(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("base_q_idx", C.c_uint32), ("bit_depth", C.c_uint32),
                ("chroma_format", C.c_uint32), ("monochrome", C.c_uint32), ("n_blocks", C.c_uint32), ("n_txb", C.c_uint32),
                ("n_coeffs", C.c_uint64), ("blocks", C.c_void_p), ("block_txb", C.c_void_p), ("txb", C.c_void_p),
                ("coeffs", C.c_void_p), ("recon", C.c_void_p * 3), ("recon_pitch", C.c_int * 3), ("aw", C.c_int),
                ("ah", C.c_int), ("tile_cols", C.c_int), ("tile_rows", C.c_int), ("fp", _FrameParams)]


class _Times(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("setup", "analysis", "dp", "recon", "pack", "download", "total")] + \
               [(n, C.c_uint32) for n in ("steps", "sbs", "blocks", "skipBlocks", "cflBlocks", "rangeFixes")] + \
               [("deviceBytes", C.c_size_t), ("schedule", C.c_uint32), ("queueBlocks", C.c_uint32), ("anStreams", C.c_uint32),
                ("graph", C.c_uint32), ("tuned", C.c_uint32), ("flopsAnalysis", C.c_uint64), ("flopsRecon", C.c_uint64),
                ("postfilter", C.c_float), ("queueOcc", C.c_uint32)]


_lib = None


def lib():
    """The HIP library (loaded once).  Raises if it is not built: there is no other compute path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libav1mi.so is not built (run __graft_entry__.build() or make -f av1gpu_amd/csrc/Makefile)")
        L = C.CDLL(LIB_PATH)
        L.av1mi_default_params.argtypes = [C.POINTER(_Params)]
        L.av1mi_check_partitions.argtypes = [C.POINTER(_Params), C.c_char_p, C.c_int]
        L.av1mi_create.argtypes = [C.POINTER(_Config), C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        L.av1mi_destroy.argtypes = [C.c_void_p]
        L.av1mi_get_config.argtypes = [C.c_void_p, C.POINTER(_Config)]
        L.av1mi_encode.argtypes = [C.c_void_p, C.POINTER(_FrameIn), C.c_int, C.POINTER(_FrameOut), C.POINTER(_Times),
                                   C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        L.av1mi_frame_out_free.argtypes = [C.POINTER(_FrameOut)]
        L.av1mi_dev_alloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.av1mi_dev_free.argtypes = [C.c_void_p]
        L.av1mi_copy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.av1mi_dev_alloc_async.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_void_p]
        L.av1mi_dev_free_async.argtypes = [C.c_void_p, C.c_void_p]
        L.av1mi_copy_h2d_async.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        L.av1mi_stream_create.argtypes = [C.POINTER(C.c_void_p)]
        L.av1mi_stream_destroy.argtypes = [C.c_void_p]
        L.av1mi_stream_sync.argtypes = [C.c_void_p]
        L.av1mi_frame_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.av1mi_frame_bytes.restype = C.c_size_t
        L.av1mi_compute_units.argtypes = []
        for fn in (L.av1mi_write_obus, L.av1mi_write_avif):
            fn.argtypes = [C.POINTER(_FrameOut), C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.c_char_p, C.c_int]
        L.av1mi_buffer_free.argtypes = [C.c_void_p]
        L.av1mi_synth_image.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_int)]
        L.av1mi_set_search_export.argtypes = [C.c_void_p, C.c_int]
        L.av1mi_set_work_count.argtypes = [C.c_void_p, C.c_int]
        L.av1mi_set_ext_candidates.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        L.av1mi_set_block_lists.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.av1mi_copy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.av1mi_blockres_export.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t]
        L.av1mi_blockres_export.restype = C.c_size_t
        L.av1mi_sbj_export.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
        L.av1mi_sbj_export.restype = C.c_size_t
        L.av1mi_leaves_export.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t]
        L.av1mi_leaves_export.restype = C.c_size_t
        L.av1mi_fwd_matrix.argtypes = [C.c_int, C.c_int, C.c_void_p]
        L.av1mi_misc_tables.argtypes = [C.c_void_p] * 7
        L.av1mi_cost_table.argtypes = [C.c_int, C.c_void_p]
        _lib = L
    return _lib


@dataclass
class GpParams:
    """Search knobs (the reference's GpParams, encoder/src/gp.h:29-68); defaults = the plain maximum-effort search."""
    lambdaMul: float = 1.0
    satdLambda: float = 1.0
    rnd: float = 0.375
    zbin: float = 0.656
    rndC: float = 0.375
    zbinC: float = 0.656
    K: int = 6
    Krecheck: int = 6
    txSearch: int = 1
    txSearchCL: int = 1
    txAllCL: int = 1
    deltas: int = 1
    uvDeltas: int = 1
    cfl: int = 1
    uvK: int = 6
    shapeMask: int = 0x3F1FFF
    partMask8: int = 0xF
    partMaskN: int = 0x3FF
    maxPart: int = 64
    uvOverheadBits: float = 1.5
    splitBias: float = 1.0
    lambdaCMul: float = 1.0

    def _c(self):
        p = _Params()
        for n in _PARAM_FLOATS + _PARAM_INTS + _PARAM_FLOATS2:
            setattr(p, n, getattr(self, n))
        return p


@dataclass
class GpConfig:
    """GpConfig (encoder/src/pipeline.h:44-105) subset for the intra RD core."""
    qindex: int = 120
    edgeFilter: int = 1
    P: GpParams = field(default_factory=GpParams)
    schedule: int = 0      # closed loop: 0 persistent queue, 1 one launch per wavefront step (same results)
    queueBlocks: int = 0   # persistent workgroups, 0 = every resident slot
    anStreams: int = 1     # analysis shape launches spread over 1..4 streams
    graph: int = 0         # capture the device part per batch shape as a hipGraph and replay it
    tune: int = 0          # the first 6 calls of a batch shape try anStreams 1/4 x closed loop (queue at 1 or 2 WG/CU, steps)
    # post-filters (pipeline.h:60-66): lfMode 0 off / 1 level from each frame's qindex (times lfMul) / 2 fixed lfLevel;
    # cdef 1 = one strength set (damping cdefDampingMinus3 + 3, primary 0..15, secondary 0/1/2/4)
    lfMode: int = 0
    lfMul: float = 1.0
    lfLevel: int = 0
    cdef: int = 0
    cdefDampingMinus3: int = 0
    cdefYPri: int = 0
    cdefYSec: int = 0
    cdefUVPri: int = 0
    cdefUVSec: int = 0
    queueOcc: int = 0      # closed-loop workgroups per CU: 0 by the batch's parallelism, 1 or 2 (same results)


@dataclass
class GpFrameIn:
    """One input frame (pipeline.h:149-163): planes padded by edge replication to aw x ah (multiples of 64).
    Planes are device pointers (ints, e.g. torch tensors' data_ptr()) or host numpy arrays (uploaded by encode)."""
    W: int
    H: int
    planes: tuple = ()
    spitch: tuple = ()
    aw: int = 0
    ah: int = 0
    mono: int = 0
    yuv444: int = 0
    bitDepth: int = 8
    qindex: int = 0


@dataclass
class GpFrameOut:
    """DecisionRecord + reconstruction of one frame (pipeline.h:165-178)."""
    width: int
    height: int
    base_q_idx: int
    bit_depth: int
    blocks: np.ndarray
    block_txb: np.ndarray
    txb: np.ndarray
    coeffs: np.ndarray
    recon: list
    tile_cols: int
    tile_rows: int
    chroma_format: int = 0       # 0 = 4:2:0 (or monochrome), 1 = 4:4:4
    fp: object = None            # the frame's decisions (av1mi_frame_params: loop filter levels, CDEF, ...)


@dataclass
class GpTimes:
    setup: float = 0
    analysis: float = 0
    dp: float = 0
    recon: float = 0
    pack: float = 0
    download: float = 0
    total: float = 0
    steps: int = 0
    sbs: int = 0
    blocks: int = 0
    skipBlocks: int = 0
    cflBlocks: int = 0
    rangeFixes: int = 0
    deviceBytes: int = 0
    schedule: int = 0
    queueBlocks: int = 0
    anStreams: int = 0
    graph: int = 0
    tuned: int = 0
    flopsAnalysis: int = 0
    flopsRecon: int = 0
    postfilter: float = 0
    queueOcc: int = 0


def check_partitions(P: GpParams):
    """gp_check_partitions (pipeline.h:216): raises ValueError for options without a finite partition everywhere."""
    err = C.create_string_buffer(256)
    p = P._c()
    if lib().av1mi_check_partitions(C.byref(p), err, 256) != 0:
        raise ValueError(err.value.decode())


def pad_planes(planes, W, H, ssx=1):
    """Edge-replicate visible planes to the 64-aligned coded area (the layout GpFrameIn expects)."""
    aw, ah = ((W + 63) // 64) * 64, ((H + 63) // 64) * 64
    out = []
    for p, a in enumerate(planes):
        sx = ssx if p else 0
        vw, vh = (W + sx) >> sx, (H + sx) >> sx
        a = a[:vh, :vw]
        out.append(np.ascontiguousarray(np.pad(a, ((0, (ah >> sx) - vh), (0, (aw >> sx) - vw)), mode="edge")))
    return out, aw, ah


def synth_image(seed, W, H, ssx=1, bd=8):
    """The benchmark's seeded synthetic image (visible planes, host numpy)."""
    dt = np.uint8 if bd == 8 else np.uint16
    cw, ch = (W + ssx) >> ssx, (H + ssx) >> ssx
    y, u, v = np.zeros((H, W), dt), np.zeros((ch, cw), dt), np.zeros((ch, cw), dt)
    st = (C.c_int * 3)(W, cw, cw)
    rc = lib().av1mi_synth_image(seed, W, H, ssx, bd, y.ctypes.data, u.ctypes.data, v.ctypes.data, st)
    if rc != 0:
        raise ValueError("synth_image: bad arguments")
    return [y, u, v]


def write_bitstream(rec, edge_filter=1, avif=True, monochrome=0):
    """The AV1 bitstream of a record (GpFrameOut, or a dict with width, height, base_q_idx, bit_depth, blocks,
    block_txb, txb, coeffs, tile_cols, tile_rows, and chroma_format 0 = 4:2:0 / 1 = 4:4:4, default 0): an AVIF file
    (avif=True) or the bare OBUs.  Host only."""
    g = (lambda k: rec[k]) if isinstance(rec, dict) else (lambda k: getattr(rec, k))
    cf = rec.get("chroma_format", 0) if isinstance(rec, dict) else getattr(rec, "chroma_format", 0)
    blocks = np.ascontiguousarray(g("blocks"))
    bt = np.ascontiguousarray(g("block_txb"), dtype=np.uint32)
    txb = np.ascontiguousarray(g("txb"))
    co = np.ascontiguousarray(g("coeffs"), dtype=np.int16)
    o = _FrameOut()
    o.width, o.height, o.base_q_idx, o.bit_depth = g("width"), g("height"), g("base_q_idx"), g("bit_depth")
    o.chroma_format, o.monochrome = cf, monochrome
    o.n_blocks, o.n_txb, o.n_coeffs = len(blocks), len(txb), len(co)
    o.blocks, o.block_txb, o.txb, o.coeffs = blocks.ctypes.data, bt.ctypes.data, txb.ctypes.data, co.ctypes.data
    o.tile_cols, o.tile_rows = g("tile_cols"), g("tile_rows")
    fp = rec.get("fp") if isinstance(rec, dict) else getattr(rec, "fp", None)
    if fp is not None:
        C.memmove(C.byref(o.fp), C.byref(fp), C.sizeof(_FrameParams))
    data, size = C.c_void_p(), C.c_size_t()
    err = C.create_string_buffer(512)
    fn = lib().av1mi_write_avif if avif else lib().av1mi_write_obus
    if fn(C.byref(o), edge_filter, C.byref(data), C.byref(size), err, 512) != 0:
        raise ValueError("write_bitstream: " + err.value.decode())
    try:
        return C.string_at(data, size.value)
    finally:
        lib().av1mi_buffer_free(data)


def device_compute_units():
    """Compute units of the current HIP device (0 without one)."""
    return int(lib().av1mi_compute_units())


def batch_spec(n=256, seed=1000):
    """The mixed-size batch workload: n images with sizes log-uniform over 0.18-24 MP, aspect ratio (W / H) uniform over
    [0.5, 2], sides rounded to multiples of 8, qindex drawn from {80, 120, 160}; image i is the synthetic image of seed
    seed + i.  Returns [(W, H, qindex, seed)]."""
    import math
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        mp = math.exp(rng.uniform(math.log(0.18), math.log(24.0)))
        a = rng.uniform(0.5, 2.0)
        W = max(8, int(round(math.sqrt(mp * 1e6 * a) / 8.0)) * 8)
        H = max(8, int(round(math.sqrt(mp * 1e6 / a) / 8.0)) * 8)
        out.append((W, H, int(rng.choice([80, 120, 160])), seed + i))
    return out


class DeviceFrame:
    """Padded planes of one frame resident in device memory (owned; freed on close / garbage collection)."""

    def __init__(self, planes, W, H, bd=8, qindex=0, stream=None, yuv444=False):
        """stream: upload with stream-ordered allocation and copies on that stream (freed on it too); yuv444: the chroma
        planes are at full resolution."""
        padded, aw, ah = pad_planes(planes, W, H, ssx=0 if yuv444 else 1)
        self.W, self.H, self.aw, self.ah, self.bd, self.qindex = W, H, aw, ah, bd, qindex
        self.yuv444 = 1 if yuv444 else 0
        self.ptrs, self.pitch, self.stream = [], [], stream
        self._host = padded if stream else None   # pageable sources must outlive the asynchronous copies
        for a in padded:
            p = C.c_void_p()
            ok = (lib().av1mi_dev_alloc_async(C.byref(p), a.nbytes, stream) if stream else lib().av1mi_dev_alloc(C.byref(p), a.nbytes)) == 0
            if not ok:
                self.close()
                raise MemoryError("device allocation of %d bytes failed" % a.nbytes)
            self.ptrs.append(p.value)
            ok = (lib().av1mi_copy_h2d_async(p, a.ctypes.data, a.nbytes, stream) if stream else lib().av1mi_copy_h2d(p, a.ctypes.data, a.nbytes)) == 0
            if not ok:
                self.close()
                raise RuntimeError("host-to-device copy failed")
            self.pitch.append(a.shape[1])

    def frame_in(self):
        return GpFrameIn(W=self.W, H=self.H, planes=tuple(self.ptrs), spitch=tuple(self.pitch), aw=self.aw, ah=self.ah,
                         bitDepth=self.bd, qindex=self.qindex, yuv444=self.yuv444)

    def close(self):
        for p in self.ptrs:
            if self.stream:
                lib().av1mi_dev_free_async(p, self.stream)
            else:
                lib().av1mi_dev_free(p)
        self.ptrs = []
        self._host = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _OutOwner:
    """Owns the host buffers of one av1mi_frame_out (the library's allocations) until garbage collection."""

    def __init__(self, o):
        self.o = _FrameOut()
        C.memmove(C.byref(self.o), C.byref(o), C.sizeof(_FrameOut))

    def __del__(self):
        try:
            lib().av1mi_frame_out_free(C.byref(self.o))
        except Exception:
            pass


def _owned(ptr, n, dtype, owner):
    """A NumPy array over n items at ptr (no copy) that keeps owner alive."""
    dt = np.dtype(dtype)
    if not n or not ptr:
        return np.zeros(0, dt)
    raw = (C.c_char * (n * dt.itemsize)).from_address(ptr)
    raw._owner = owner
    return np.frombuffer(raw, dt)


class GpEncoder:
    """GpEncoder (encoder/src/pipeline.h:191-211): init(config) once, encode(frames) per batch."""

    def __init__(self):
        self._h = None
        self._cfg = None

    def init(self, cfg: GpConfig):
        if self._h:
            lib().av1mi_destroy(self._h)
            self._h = None
        c = _Config(cfg.qindex, cfg.edgeFilter, cfg.P._c(), cfg.schedule, cfg.queueBlocks, cfg.anStreams, cfg.graph, cfg.tune,
                    cfg.lfMode, cfg.lfMul, cfg.lfLevel, *[getattr(cfg, n) for n in _POSTFILTER_INTS], cfg.queueOcc)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        if lib().av1mi_create(C.byref(c), C.byref(h), err, 512) != 0:
            raise RuntimeError("GpEncoder.init: " + err.value.decode())
        self._h = h.value
        self._cfg = cfg
        return self

    def config(self) -> GpConfig:
        return self._cfg

    def encode(self, frames, stream=None, want_recon=True):
        """Encode a batch; frames: GpFrameIn with device pointers, or DeviceFrame.  Returns ([GpFrameOut], GpTimes)."""
        if not self._h:
            raise RuntimeError("GpEncoder.encode before init")
        fins = [f.frame_in() if isinstance(f, DeviceFrame) else f for f in frames]
        n = len(fins)
        arr = (_FrameIn * max(n, 1))()
        for i, f in enumerate(fins):
            a = arr[i]
            a.W, a.H, a.mono, a.yuv444, a.bitDepth = f.W, f.H, f.mono, f.yuv444, f.bitDepth
            for p in range(3):
                a.dsrc[p] = f.planes[p]
                a.spitch[p] = f.spitch[p]
            a.aw, a.ah, a.qindex = f.aw, f.ah, f.qindex
        outs = (_FrameOut * max(n, 1))()
        t = _Times()
        err = C.create_string_buffer(512)
        rc = lib().av1mi_encode(self._h, arr, n, outs, C.byref(t), stream or None, 1 if want_recon else 0, err, 512)
        if rc != 0:
            raise RuntimeError("GpEncoder.encode: " + err.value.decode())
        res = []
        try:
            for i in range(n):
                o = outs[i]
                own = _OutOwner(o)          # frees the host buffers when the last array over them is gone
                o = own.o
                outs[i] = _FrameOut()       # ownership moved to `own`
                nb, nt, nc = o.n_blocks, o.n_txb, o.n_coeffs
                blocks = _owned(o.blocks, nb, BLOCK_DTYPE, own)
                bt = _owned(o.block_txb, nb + 1, np.uint32, own)
                txb = _owned(o.txb, nt, TXB_DTYPE, own)
                co = _owned(o.coeffs, nc, np.int16, own)
                rec = []
                if want_recon:
                    bps = 2 if o.bit_depth > 8 else 1
                    for p in range(3):
                        h = o.ah >> (1 if p and o.chroma_format != 1 else 0)
                        w = o.recon_pitch[p]
                        rec.append(_owned(o.recon[p], w * h, np.uint16 if bps == 2 else np.uint8, own).reshape(h, w))
                fp = _FrameParams()
                C.memmove(C.byref(fp), C.byref(o.fp), C.sizeof(_FrameParams))
                res.append(GpFrameOut(o.width, o.height, o.base_q_idx, o.bit_depth, blocks, bt, txb, co, rec, o.tile_cols,
                                      o.tile_rows, o.chroma_format, fp))
        finally:
            for i in range(n):
                lib().av1mi_frame_out_free(C.byref(outs[i]))
        times = GpTimes(**{k: getattr(t, k) for k, _ in _Times._fields_})
        return res, times

    def set_ext_candidates(self, cand, off, topk):
        """External screener (GpBatch::extCand / extOff / extTopK): cand {shape: device pointer of av1mi_ext_cand
        entries}, off: device pointer of 22 * n int64 offsets (n frames per call), topk entries per block; None clears."""
        arr = (C.c_void_p * 22)()
        for s, p in (cand or {}).items():
            arr[s] = p
        err = C.create_string_buffer(512)
        if lib().av1mi_set_ext_candidates(self._h, arr if cand else None, off, topk, err, 512) != 0:
            raise ValueError("set_ext_candidates: " + err.value.decode())

    def set_block_lists(self, lists, counts):
        """Block lists (GpAnX): lists / counts {shape: device pointer}; None clears."""
        la, ca = (C.c_void_p * 22)(), (C.c_void_p * 22)()
        for s, p in (lists or {}).items():
            la[s] = p
            ca[s] = counts[s]
        err = C.create_string_buffer(512)
        if lib().av1mi_set_block_lists(self._h, la if lists else None, ca if lists else None, err, 512) != 0:
            raise ValueError("set_block_lists: " + err.value.decode())

    def set_work_count(self, on=True):
        """Count the forward-transform FLOPs of each later encode (GpTimes.flopsAnalysis / flopsRecon)."""
        lib().av1mi_set_work_count(self._h, 1 if on else 0)

    def set_search_export(self, on=True):
        """Keep the search internals of each later encode (analysis records, superblock costs, leaves)."""
        lib().av1mi_set_search_export(self._h, 1 if on else 0)

    def search_export(self, img=0):
        """The last encode's search internals of frame img: {"blockres": {shape: records}, "sbJ", "leaf_count",
        "leaves" [nSb, 256]}."""
        L, h = lib(), self._h
        br = {}
        for s in SHAPES:
            n = int(L.av1mi_blockres_export(h, img, s, None, 0))
            a = np.zeros(n, BLOCKRES_DTYPE)
            L.av1mi_blockres_export(h, img, s, a.ctypes.data, n)
            br[s] = a
        nsb = int(L.av1mi_sbj_export(h, img, None, 0))
        sbj = np.zeros(nsb, np.float32)
        L.av1mi_sbj_export(h, img, sbj.ctypes.data, nsb)
        cnt = np.zeros(nsb, np.uint16)
        lv = np.zeros(nsb * 256, LEAF_DTYPE)
        L.av1mi_leaves_export(h, img, cnt.ctypes.data, lv.ctypes.data, nsb)
        return {"blockres": br, "sbJ": sbj, "leaf_count": cnt, "leaves": lv.reshape(nsb, 256)}

    def close(self):
        if self._h:
            lib().av1mi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
