"""av1mi_encode_noalloc (the entry point that allocates no device memory: its buffers live in a file-scope scratch array,
the records go into the caller's buffers) against av1mi_encode_dev, the existing entry point it varies, through the C
ABI: bit-identical records (blocks, each block's first tx block, tx blocks, coefficients), counts and reconstruction,
on two calls in a row, the second with a smaller batch (the scratch array keeps the first call's contents); the learned
pruning path likewise; batches beyond the scratch array or the output capacities are refused with a message."""
import ctypes as C

import numpy as np
import pytest

import av1gpu_amd as A
from av1gpu_amd import learn

pytestmark = pytest.mark.gpu


class _DevRecords(C.Structure):
    _fields_ = [("n_frames", C.c_uint32), ("blk0", C.c_void_p), ("txb0", C.c_void_p), ("coef0", C.c_void_p),
                ("blocks", C.c_void_p), ("block_txb_rel", C.c_void_p), ("txb", C.c_void_p), ("coeffs", C.c_void_p),
                ("pool", C.c_void_p)]


class _NoAllocOut(C.Structure):
    _fields_ = [("blocks", C.c_void_p), ("block_txb", C.c_void_p), ("txb", C.c_void_p), ("coeffs", C.c_void_p),
                ("cap_blocks", C.c_uint64), ("cap_txb", C.c_uint64), ("cap_coeffs", C.c_uint64),
                ("recon", C.POINTER(C.c_void_p))]


def _lib():
    L = A.lib()
    L.av1mi_encode_dev.argtypes = [C.c_void_p, C.POINTER(A._FrameIn), C.c_int, C.POINTER(A._FrameOut), C.POINTER(A._Times),
                                   C.c_void_p, C.c_int, C.POINTER(_DevRecords), C.c_char_p, C.c_int]
    L.av1mi_dev_records_free.argtypes = [C.POINTER(_DevRecords), C.c_void_p]
    L.av1mi_encode_noalloc.argtypes = [C.c_void_p, C.POINTER(A._FrameIn), C.c_int, C.POINTER(_NoAllocOut),
                                       C.POINTER(A._FrameOut), C.POINTER(A._Times), C.c_void_p, C.c_char_p, C.c_int]
    L.av1mi_copy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    return L


def _frames_in(frames):
    arr = (A._FrameIn * len(frames))()
    for i, f in enumerate(frames):
        a, g = arr[i], f.frame_in()
        a.W, a.H, a.mono, a.yuv444, a.bitDepth = g.W, g.H, g.mono, g.yuv444, g.bitDepth
        for p in range(3):
            a.dsrc[p] = g.planes[p]
            a.spitch[p] = g.spitch[p]
        a.aw, a.ah, a.qindex = g.aw, g.ah, g.qindex
    return arr


def _d2h(L, ptr, n, dtype):
    a = np.zeros(n, dtype)
    if n:
        assert L.av1mi_copy_d2h(a.ctypes.data, ptr, a.nbytes) == 0
    return a


def _planes_shape(o, p):
    ss = 0 if o.chroma_format == 1 else 1
    return (o.ah >> (ss if p else 0), o.aw >> (ss if p else 0))


def _encode_dev(L, enc, frames):
    n = len(frames)
    outs, t, dev = (A._FrameOut * n)(), A._Times(), _DevRecords()
    err = C.create_string_buffer(512)
    assert L.av1mi_encode_dev(enc._h, _frames_in(frames), n, outs, C.byref(t), None, 1, C.byref(dev), err, 512) == 0, err.value
    nb, nt = (C.c_uint32 * (n + 1)).from_address(dev.blk0), (C.c_uint32 * (n + 1)).from_address(dev.txb0)
    nc = (C.c_uint64 * (n + 1)).from_address(dev.coef0)
    rec = {"blocks": _d2h(L, dev.blocks, nb[n], A.BLOCK_DTYPE), "block_txb": _d2h(L, dev.block_txb_rel, nb[n], np.uint32),
           "txb": _d2h(L, dev.txb, nt[n], A.TXB_DTYPE), "coeffs": _d2h(L, dev.coeffs, nc[n], np.int16),
           "counts": [(outs[i].n_blocks, outs[i].n_txb, outs[i].n_coeffs) for i in range(n)], "recon": []}
    for i in range(n):
        o = outs[i]
        dt = np.uint16 if o.bit_depth > 8 else np.uint8
        rec["recon"].append([np.ctypeslib.as_array((C.c_uint8 * (int(np.prod(_planes_shape(o, p))) * np.dtype(dt).itemsize))
                                                   .from_address(o.recon[p])).view(dt).reshape(_planes_shape(o, p)).copy()
                             for p in range(3)])
        L.av1mi_frame_out_free(C.byref(outs[i]))
    L.av1mi_dev_records_free(C.byref(dev), None)
    return rec


class _Buffers:
    """Device output buffers of the no-allocation entry point (allocated here, by the test), with room to spare."""

    def __init__(self, L, caps, frames):
        self.L, self.ptrs = L, []
        self.o = _NoAllocOut()
        self.caps = caps
        for name, cap, size in (("blocks", caps[0], 16), ("block_txb", caps[0], 4), ("txb", caps[1], 16), ("coeffs", caps[2], 2)):
            setattr(self.o, name, self._alloc(cap * size))
        self.o.cap_blocks, self.o.cap_txb, self.o.cap_coeffs = caps
        self.recon = (C.c_void_p * (3 * len(frames)))()
        for i, f in enumerate(frames):
            ss = 0 if f.yuv444 else 1
            bps = 2 if f.bd > 8 else 1
            for p in range(3):
                self.recon[3 * i + p] = self._alloc((f.aw >> (ss if p else 0)) * (f.ah >> (ss if p else 0)) * bps)
        self.o.recon = C.cast(self.recon, C.POINTER(C.c_void_p))

    def _alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.av1mi_dev_alloc(C.byref(p), max(nbytes, 256)) == 0
        self.ptrs.append(p.value)
        return p.value

    def close(self):
        for p in self.ptrs:
            self.L.av1mi_dev_free(p)
        self.ptrs = []


def _encode_noalloc(L, enc, frames, buf):
    n = len(frames)
    outs, t = (A._FrameOut * n)(), A._Times()
    err = C.create_string_buffer(512)
    rc = L.av1mi_encode_noalloc(enc._h, _frames_in(frames), n, C.byref(buf.o), outs, C.byref(t), None, err, 512)
    if rc != 0:
        return None, err.value.decode()
    counts = [(outs[i].n_blocks, outs[i].n_txb, outs[i].n_coeffs) for i in range(n)]
    tb, tt, tc = (sum(c[k] for c in counts) for k in range(3))
    rec = {"blocks": _d2h(L, buf.o.blocks, tb, A.BLOCK_DTYPE), "block_txb": _d2h(L, buf.o.block_txb, tb, np.uint32),
           "txb": _d2h(L, buf.o.txb, tt, A.TXB_DTYPE), "coeffs": _d2h(L, buf.o.coeffs, tc, np.int16), "counts": counts,
           "recon": []}
    for i in range(n):
        o = outs[i]
        assert not o.blocks and not o.txb and not o.coeffs and not o.recon[0]
        dt = np.uint16 if o.bit_depth > 8 else np.uint8
        rec["recon"].append([_d2h(L, buf.recon[3 * i + p], int(np.prod(_planes_shape(o, p))), dt).reshape(_planes_shape(o, p))
                             for p in range(3)])
    return rec, ""


def _same(a, b):
    assert a["counts"] == b["counts"]
    for k in ("blocks", "block_txb", "txb", "coeffs"):
        assert len(a[k]) == len(b[k]) and np.array_equal(a[k], b[k]), k
    for ra, rb in zip(a["recon"], b["recon"]):
        for p in range(3):
            assert np.array_equal(ra[p], rb[p]), "recon plane %d" % p


def _batch(specs, bd=8, yuv444=False):
    return [A.DeviceFrame(A.synth_image(s, W, H, ssx=0 if yuv444 else 1, bd=bd), W, H, bd=bd, qindex=q, yuv444=yuv444)
            for W, H, s, q in specs]


@pytest.mark.parametrize("learned", [False, True])
def test_noalloc_equals_encode_dev_on_two_calls(learned):
    L = _lib()
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
    if learned:
        learn.set_shortlist(enc, learn.load_params(), 0.8)
    big = _batch([(1000, 712, 7001, 120), (328, 200, 7002, 80), (136, 72, 7003, 160)])
    small = _batch([(200, 136, 7004, 120)])
    bufs = []
    try:
        for k, frames in enumerate((big, small)):
            print("noalloc call %d (learned=%s): %s" % (k, learned, [(f.W, f.H, f.qindex) for f in frames]), flush=True)
            ref = _encode_dev(L, enc, frames)
            caps = tuple(sum(c[j] for c in ref["counts"]) + 64 for j in range(3))
            buf = _Buffers(L, caps, frames)
            bufs.append(buf)
            got, err = _encode_noalloc(L, enc, frames, buf)
            assert got is not None, err
            _same(ref, got)
    finally:
        for b in bufs:
            b.close()
        for f in big + small:
            f.close()
        enc.close()


def test_noalloc_refuses_what_does_not_fit():
    L = _lib()
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
    frames = _batch([(264, 200, 7005, 120)])
    try:
        ref = _encode_dev(L, enc, frames)
        nb = ref["counts"][0][0]
        buf = _Buffers(L, (nb - 1, 10 ** 6, 10 ** 7), frames)           # one block short
        got, err = _encode_noalloc(L, enc, frames, buf)
        assert got is None and "output buffers hold" in err
        buf.close()
        buf = _Buffers(L, (nb, 10 ** 6, 10 ** 7), frames)               # then exactly enough: the same records
        got, err = _encode_noalloc(L, enc, frames, buf)
        assert got is not None, err
        _same(ref, got)
        buf.close()
    finally:
        for f in frames:
            f.close()
    # a frame whose buffers exceed the scratch array: refused before any launch (its planes are never read)
    W, H = 8192, 8192
    planes = []
    for w, h in ((W, H), (W // 2, H // 2), (W // 2, H // 2)):
        p = C.c_void_p()
        assert L.av1mi_dev_alloc(C.byref(p), w * h) == 0
        planes.append(p.value)
    fin = (A._FrameIn * 1)()
    fin[0].W, fin[0].H, fin[0].bitDepth, fin[0].aw, fin[0].ah = W, H, 8, W, H
    for p in range(3):
        fin[0].dsrc[p] = planes[p]
        fin[0].spitch[p] = W if p == 0 else W // 2
    buf = _Buffers(L, (10, 10, 10), [])
    outs, t = (A._FrameOut * 1)(), A._Times()
    err = C.create_string_buffer(512)
    assert L.av1mi_encode_noalloc(enc._h, fin, 1, C.byref(buf.o), outs, C.byref(t), None, err, 512) == -1
    assert b"device scratch" in err.value
    buf.close()
    for p in planes:
        L.av1mi_dev_free(p)
    enc.close()
