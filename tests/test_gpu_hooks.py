"""The analysis hooks on the GPU, through the C ABI, against the oracle on the same frames: the external screener
(av1mi_set_ext_candidates: the reference's GpBatch::extCand / extOff / extTopK, encoder/src/gp.h:136-143 and 164-171)
with random candidates, unusable entries and out-of-range deltas over a two-frame batch (per-frame offsets), and block
lists (av1mi_set_block_lists: GpAnX, encoder/src/an_stage.h:23-27) with random subsets; records, reconstruction and the
analysis records equal the oracle's.  Lists that would leave a superblock no partition, and lists together with the
learned pruning, are refused."""
import ctypes as C

import numpy as np
import pytest

import av1gpu_amd as A
from av1gpu_amd import learn
from refcheck import EXT_DTYPE
from test_gpu_parity import same_search
from test_oracle_hooks import SHAPES, grid, same_record

pytestmark = pytest.mark.gpu

FRAMES = [(200, 136, 3, 160), (136, 72, 50, 120)]


class _Dev:
    def __init__(self):
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert A.lib().av1mi_dev_alloc(C.byref(p), max(a.nbytes, 16)) == 0
        if a.nbytes:
            assert A.lib().av1mi_copy_h2d(p, a.ctypes.data, a.nbytes) == 0
        self.ptrs.append(p.value)
        return p.value

    def close(self):
        for p in self.ptrs:
            A.lib().av1mi_dev_free(p)
        self.ptrs = []


def _gpu_vs_oracle(oracle, enc_setup, per_frame_hooks):
    planes = [A.synth_image(s, W, H) for W, H, s, q in FRAMES]
    frames = [A.DeviceFrame(p, W, H, qindex=q) for p, (W, H, s, q) in zip(planes, FRAMES)]
    dev = _Dev()
    e = A.GpEncoder().init(A.GpConfig(qindex=120))
    try:
        enc_setup(e, dev)
        e.set_search_export(True)
        g, _ = e.encode(frames)
        for i, (p, (W, H, s, q)) in enumerate(zip(planes, FRAMES)):
            r = oracle.encode(p, W, H, qindex=q, threads=4, detail=True, **per_frame_hooks(i))
            same_record(dict(blocks=g[i].blocks, block_txb=g[i].block_txb, txb=g[i].txb, coeffs=g[i].coeffs,
                             recon=[g[i].recon[k].astype(np.uint16) for k in range(3)]), r)
            same_search(e.search_export(i), r)
    finally:
        e.close()
        dev.close()
        for f in frames:
            f.close()


def test_gpu_ext_candidates_equal_oracle(oracle):
    rng = np.random.default_rng(7)
    topk = 5
    ext = {}   # per frame, per shape: [blocks, topk]
    for i, (W, H, s, q) in enumerate(FRAMES):
        ext[i] = {}
        for sh in SHAPES:
            nx, ny = grid(sh, W, H)
            a = np.zeros((nx * ny, topk), EXT_DTYPE)
            a["mode"] = rng.integers(0, 17, size=a.shape)          # 13..16: skipped
            a["delta"] = rng.integers(-4, 5, size=a.shape)         # -4, 4: outside the range, taken as 0
            a["J"] = rng.random(a.shape).astype(np.float32)
            a["mode"][rng.random(nx * ny) < 0.1] = 99              # some blocks without a usable entry: screened
            ext[i][sh] = a
    shapes = [sh for sh in SHAPES if sh != 12]                     # 64x64 keeps its screening

    def setup(e, dev):
        cand, off = {}, np.zeros(22 * len(FRAMES), np.int64)
        for sh in shapes:
            base = 0
            for i in range(len(FRAMES)):
                off[sh * len(FRAMES) + i] = base
                base += len(ext[i][sh])
            cand[sh] = dev.put(np.concatenate([ext[i][sh] for i in range(len(FRAMES))]))
        e.set_ext_candidates(cand, dev.put(off), topk)

    _gpu_vs_oracle(oracle, setup, lambda i: dict(ext={sh: ext[i][sh] for sh in shapes}, ext_topk=topk))


def test_gpu_block_lists_equal_oracle(oracle):
    rng = np.random.default_rng(11)
    shapes = (1, 2, 4, 5, 16, 17, 18, 19, 20, 21)                  # 2:1 and 4:1 shapes: squares keep every SB partitionable
    masks = {}
    for i, (W, H, s, q) in enumerate(FRAMES):
        masks[i] = {sh: (rng.random(np.prod(grid(sh, W, H))) < 0.5).astype(np.uint8) for sh in shapes}

    def setup(e, dev):
        lists, counts = {}, {}
        for sh in shapes:
            idx, base = [], 0
            for i in range(len(FRAMES)):
                idx.append(np.nonzero(masks[i][sh])[0] + base)
                base += len(masks[i][sh])
            li = np.concatenate(idx).astype(np.uint32)
            rng.shuffle(li)                                          # order does not matter
            li = np.concatenate([li, np.array([base + 5, 1 << 30], np.uint32)])   # entries outside the shape: ignored
            lists[sh] = dev.put(li)
            counts[sh] = dev.put(np.array([len(li)], np.uint32))
        e.set_block_lists(lists, counts)

    _gpu_vs_oracle(oracle, setup, lambda i: dict(analyse=masks[i]))


def test_block_lists_refused_where_they_could_leave_no_partition():
    dev = _Dev()
    e = A.GpEncoder().init(A.GpConfig(qindex=120))
    try:
        p = dev.put(np.zeros(4, np.uint32))
        with pytest.raises(ValueError, match="partition"):
            e.set_block_lists({s: p for s in SHAPES}, {s: p for s in SHAPES})
        learn.set_shortlist(e, learn.load_params(), 0.8)
        with pytest.raises(ValueError, match="learned"):
            e.set_block_lists({17: p}, {17: p})
    finally:
        e.close()
        dev.close()
