"""Learned candidate pruning: labels, training, inference helpers and the checks of the network kernels.

The network (av1gpu_amd/csrc/k_learn.hip) sees the 64x64 luma of a superblock and the qindex and predicts
  * 85 "expand" outputs, one per square partition node (the 64x64 node, four 32x32, sixteen 16x16, sixty-four 8x8, raster
    order per level): 1 when some leaf of the exhaustive search inside the node is at most half its size in both
    dimensions (its children are used);
  * 85 "own" outputs, same nodes: 1 when some leaf inside the node is one of the node's N x N, N x N/2 or N/2 x N shapes;
  * 85 "four" outputs, same nodes: 1 when some leaf inside the node is one of its 4-way shapes (N x N/4, N/4 x N);
  * 13 luma-mode outputs: 1 when some block of the superblock uses that luma mode in the exhaustive result.
The encoder then analyses a node's children / own shapes / 4-way shapes unless the network is confident (probability
at most 1 - conf) that they are unused (a node that would lose all three keeps its own shapes), and screens only the 8
most likely modes where it is confident (the product of 1 - p over the 5 others at least conf) that none of the others
is used.  With cost-aware labels (rd_labels) "used" becomes "dropping them costs more than tau".

Training data are the exhaustive search's own decisions on seeded synthetic images; training runs on the GPU with the
library's forward (MFMA) and backward kernels.  tests/ln_ref.py restates the same network in float64 NumPy with the
BF16 rounding at the same points, and the decisions; the tests check the kernels' activations, gradients and the
encoder's decisions with it.
"""
import ctypes as C
import math
import os

import numpy as np

import av1gpu_amd as A

N_SPLIT, OWN0, FOUR0, MODE0, N_OUT, N_FEAT = 85, 85, 170, 255, 268, 512   # node count / first own, 4-way, mode output
LEVEL_BASE = (0, 1, 5, 21)
KP = (32, 288, 288, 288)
CIN = (2, 32, 32, 32)


def _sigs():
    L = A.lib()
    if getattr(L, "_ln_sigs", False):
        return L
    P, U8, F = C.c_void_p, C.c_void_p, C.c_void_p
    L.av1mi_ln_param_count.restype = C.c_size_t
    L.av1mi_ln_init_params.argtypes = [F, C.c_uint64]
    L.av1mi_ln_create.argtypes = [F, C.POINTER(C.c_void_p)]
    L.av1mi_ln_destroy.argtypes = [P]
    L.av1mi_ln_get_params.argtypes = [P, F]
    L.av1mi_ln_forward.argtypes = [P, U8, U8, C.c_uint32, F]
    L.av1mi_ln_activations.argtypes = [P, C.c_int, C.c_uint32, P]
    L.av1mi_ln_set_data.argtypes = [P, U8, U8, F, F, C.c_uint32]
    L.av1mi_ln_gradient.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float), F]
    L.av1mi_ln_train_step.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.POINTER(C.c_float)]
    L.av1mi_set_shortlist.argtypes = [P, F, C.c_float, C.c_char_p, C.c_int]
    L.av1mi_set_dp_export.argtypes = [P, C.c_int]
    L.av1mi_dp_export.argtypes = [P, F, C.c_size_t]
    L.av1mi_dp_export.restype = C.c_size_t
    L._ln_sigs = True
    return L


def param_count():
    return int(_sigs().av1mi_ln_param_count())


def init_params(seed=1):
    p = np.zeros(param_count(), np.float32)
    _sigs().av1mi_ln_init_params(p.ctypes.data, seed)
    return p


def unpack(p):
    """Views of the parameter vector: [(W_l [32][KP_l], b_l [32]) for l in 0..3], W5 [268][512], b5 [268]."""
    out, off = [], 0
    for l in range(4):
        w = p[off:off + 32 * KP[l]].reshape(32, KP[l])
        off += 32 * KP[l]
        b = p[off:off + 32]
        off += 32
        out.append((w, b))
    w5 = p[off:off + N_OUT * N_FEAT].reshape(N_OUT, N_FEAT)
    off += N_OUT * N_FEAT
    return out, w5, p[off:off + N_OUT]


# ---------------------------------------------------------------- data
def sb_tiles(y, W, H):
    """64x64 luma tiles of every superblock (edge-replicated padding), raster order: [nSb, 4096] uint8."""
    aw, ah = (W + 63) // 64 * 64, (H + 63) // 64 * 64
    p = np.pad(y[:H, :W], ((0, ah - H), (0, aw - W)), mode="edge")
    t = p.reshape(ah // 64, 64, aw // 64, 64).transpose(0, 2, 1, 3).reshape(-1, 4096)
    return np.ascontiguousarray(t)


def sb_labels(blocks, W, H):
    """Labels and label masks [nSb, 268] of a record's leaves (see the module docstring); nodes reaching past the
    frame are masked out (the encoder keeps all their shapes regardless)."""
    MiCols, MiRows = 2 * ((W + 7) >> 3), 2 * ((H + 7) >> 3)
    sbc, sbr = (MiCols + 15) >> 4, (MiRows + 15) >> 4
    lab = np.zeros((sbr * sbc, N_OUT), np.float32)
    msk = np.ones((sbr * sbc, N_OUT), np.float32)
    FW, FH = MiCols * 4, MiRows * 4
    bw = np.array([4 * (1 << v) for v in _MI_W_LOG2])[blocks["bsize"]]
    bh = np.array([4 * (1 << v) for v in _MI_H_LOG2])[blocks["bsize"]]
    x, y = blocks["mi_col"].astype(np.int64) * 4, blocks["mi_row"].astype(np.int64) * 4
    sb = (y // 64) * sbc + x // 64
    lx, ly = x % 64, y % 64
    m = np.maximum(bw, bh)
    four = np.maximum(bw, bh) == 4 * np.minimum(bw, bh)
    for lv in range(4):
        n = 64 >> lv
        node = LEVEL_BASE[lv] + (ly // n) * (64 // n) + lx // n
        sel = n >= 2 * m
        lab[sb[sel], node[sel]] = 1.0
        sel = (n == m) & ~four
        lab[sb[sel], OWN0 + node[sel]] = 1.0
        sel = (n == m) & four
        lab[sb[sel], FOUR0 + node[sel]] = 1.0
        per = 1 << lv
        for ny in range(per):
            for nx in range(per):
                nd = LEVEL_BASE[lv] + ny * per + nx
                sxs = (np.arange(sbr * sbc) % sbc) * 64 + nx * n
                sys = (np.arange(sbr * sbc) // sbc) * 64 + ny * n
                out = (sxs + n > FW) | (sys + n > FH)
                msk[out, nd] = 0.0
                msk[out, OWN0 + nd] = 0.0
                msk[out, FOUR0 + nd] = 0.0
    # modes: used by some block of the superblock
    lab[sb, MODE0 + blocks["y_mode"].astype(np.int64)] = 1.0
    return lab, msk


def dp_export(enc, on=True):
    """Make the encoder keep the partition DP's per-node relative costs (on) and return those of its last call:
    [nSb, 255] float32, superblocks in batch order (see av1mi_set_dp_export)."""
    L = _sigs()
    if on is not None:
        L.av1mi_set_dp_export(enc._h, 1 if on else 0)
    n = int(L.av1mi_dp_export(enc._h, None, 0))
    out = np.zeros(n, np.float32)
    if n:
        L.av1mi_dp_export(enc._h, out.ctypes.data, n)
    return out.reshape(-1, 3 * N_SPLIT)


def rd_labels(rel, lab, tau, tau_root=None):
    """Cost-aware node labels: a node needs its children (own shapes, 4-way shapes) when the DP's best cost without
    them is more than tau (relative) above its best -- tau_root for the 64x64 node, if given; the mode labels of lab
    stay.  rel: [nSb, 255] of dp_export for one image."""
    t = np.full(N_SPLIT, tau, np.float32)
    if tau_root is not None:
        t[0] = tau_root
    out = lab.copy()
    for k, o in enumerate((0, OWN0, FOUR0)):
        out[:, o:o + N_SPLIT] = (rel[:, k * N_SPLIT:(k + 1) * N_SPLIT] > t).astype(np.float32)
    return out


_MI_W_LOG2 = (0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 0, 2, 1, 3, 2, 4)
_MI_H_LOG2 = (0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4, 5, 4, 5, 2, 0, 3, 1, 4, 2)


# ---------------------------------------------------------------- device learner
class Learner:
    def __init__(self, params):
        self.L = _sigs()
        self.h = C.c_void_p()
        p = np.ascontiguousarray(params, np.float32)
        if self.L.av1mi_ln_create(p.ctypes.data, C.byref(self.h)) != 0:
            raise RuntimeError("av1mi_ln_create failed")
        self.n = 0

    def params(self):
        p = np.zeros(param_count(), np.float32)
        if self.L.av1mi_ln_get_params(self.h, p.ctypes.data) != 0:
            raise RuntimeError("av1mi_ln_get_params failed")
        return p

    def forward(self, tiles, q):
        tiles = np.ascontiguousarray(tiles, np.uint8)
        q = np.ascontiguousarray(q, np.uint8)
        out = np.zeros((len(tiles), N_OUT), np.float32)
        if self.L.av1mi_ln_forward(self.h, tiles.ctypes.data, q.ctypes.data, len(tiles), out.ctypes.data) != 0:
            raise RuntimeError("av1mi_ln_forward failed")
        return out

    def activations(self, layer, n):
        per = (64 * 64 * 2, 32 * 32 * 32, 16 * 16 * 32, 8 * 8 * 32, N_FEAT)[layer]
        a = np.zeros(n * per, np.uint16)
        if self.L.av1mi_ln_activations(self.h, layer, n, a.ctypes.data) != 0:
            raise RuntimeError("av1mi_ln_activations failed")
        return (a.astype(np.uint32) << 16).view(np.float32)

    def set_data(self, tiles, q, label, lmask):
        self._keep = [np.ascontiguousarray(tiles, np.uint8), np.ascontiguousarray(q, np.uint8),
                      np.ascontiguousarray(label, np.float32), np.ascontiguousarray(lmask, np.float32)]
        t, qq, lb, lm = self._keep
        if self.L.av1mi_ln_set_data(self.h, t.ctypes.data, qq.ctypes.data, lb.ctypes.data, lm.ctypes.data, len(t)) != 0:
            raise RuntimeError("av1mi_ln_set_data failed")
        self.n = len(t)

    def gradient(self, first, n):
        g = np.zeros(param_count(), np.float32)
        loss = C.c_float()
        if self.L.av1mi_ln_gradient(self.h, first, n, C.byref(loss), g.ctypes.data) != 0:
            raise RuntimeError("av1mi_ln_gradient failed")
        return loss.value, g

    def step(self, first, n, lr, wd=1e-5):
        loss = C.c_float()
        if self.L.av1mi_ln_train_step(self.h, first, n, lr, wd, C.byref(loss)) != 0:
            raise RuntimeError("av1mi_ln_train_step failed")
        return loss.value

    def close(self):
        if self.h:
            self.L.av1mi_ln_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shortlist_params.npy")


def load_params(path=PARAMS_FILE):
    """The trained network parameters shipped with the package (written by av1gpu_amd.train_shortlist)."""
    p = np.load(path, allow_pickle=False).astype(np.float32)
    if p.shape != (param_count(),):
        raise ValueError("%s holds %s parameters, the network has %d" % (path, p.shape, param_count()))
    return p


def set_shortlist(enc, params, confidence=0.8):
    """Enable (params) or disable (None) the learned pruning of an encoder."""
    L = _sigs()
    err = C.create_string_buffer(256)
    p = None if params is None else np.ascontiguousarray(params, np.float32)
    if L.av1mi_set_shortlist(enc._h, None if p is None else p.ctypes.data, confidence, err, 256) != 0:
        raise RuntimeError("set_shortlist: " + err.value.decode())


# ---------------------------------------------------------------- rate / distortion
def psnr_y(src, rec, W, H, bd=8):
    d = src[:H, :W].astype(np.float64) - rec[:H, :W].astype(np.float64)
    mse = float(np.mean(d * d))
    peak = (1 << bd) - 1
    return 99.0 if mse == 0 else 10.0 * math.log10(peak * peak / mse)


def rd_curve(enc, images, qs, bd=8, yuv444=False):
    """Rate-distortion points of an encoder over images [(planes, W, H)] at each qindex of qs: total bits of this
    build's own bitstream writer (bare OBUs), mean PSNR-Y, and the encode seconds (one call per qindex, the planes
    uploaded beforehand, one untimed warm-up call on the first image)."""
    import time
    bits, psnr, secs = [], [], 0.0
    for q in qs:
        frames = [A.DeviceFrame(p, W, H, bd=bd, qindex=q, yuv444=yuv444) for p, W, H in images]
        enc.encode(frames[:1])
        A.lib().av1mi_device_sync()
        t0 = time.perf_counter()
        out, _ = enc.encode(frames)
        secs += time.perf_counter() - t0
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(min(16, max(1, len(out)))) as ex:   # the writer releases the GIL (ctypes)
            sizes = list(ex.map(lambda o: len(A.write_bitstream(o, avif=False)), out))
        bits.append(sum(8 * n for n in sizes))
        psnr.append(float(np.mean([psnr_y(p[0], o.recon[0], W, H, bd) for (p, W, H), o in zip(images, out)])))
        for f in frames:
            f.close()
    return bits, psnr, secs


def bd_rate(r_a, p_a, r_b, p_b):
    """Bjontegaard delta rate (percent) of curve b against curve a: cubic fits of log10(rate) over PSNR, averaged over
    the PSNR range both curves cover (positive: b needs more bits for the same quality)."""
    la, lb = np.log10(np.asarray(r_a, float)), np.log10(np.asarray(r_b, float))
    pa, pb = np.asarray(p_a, float), np.asarray(p_b, float)
    fa, fb = np.polyfit(pa, la, 3), np.polyfit(pb, lb, 3)
    lo, hi = max(pa.min(), pb.min()), min(pa.max(), pb.max())
    ia, ib = np.polyint(fa), np.polyint(fb)
    avg = ((np.polyval(ib, hi) - np.polyval(ib, lo)) - (np.polyval(ia, hi) - np.polyval(ia, lo))) / (hi - lo)
    return (10.0 ** avg - 1.0) * 100.0
