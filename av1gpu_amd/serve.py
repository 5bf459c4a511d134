"""Serving loop over the intra RD core: continuous batching of single-image requests on one device.

Requests (host planes of one 8-bit 4:2:0 image each) arrive at any time.  A request is admitted when the device bytes
its encode will occupy (``av1mi_frame_bytes``) fit under the HBM budget next to everything already in flight.  Each
of ``workers`` host threads owns a non-blocking HIP stream and a ``GpEncoder`` whose persistent closed loop is limited
to its share of the compute units; a free worker takes every admitted request waiting (oldest first, up to
``max_batch``), uploads their planes with stream-ordered allocations, encodes them as one batch and frees them.  So
requests that arrive while others are on the device start as soon as a worker and memory are free, and their
analysis and closed-loop kernels overlap with the other workers' on the device.  An image's record does not depend
on what it was batched with (the encode is per image bit-exact), which the serving test checks.

No counterpart in the reference, whose batch front end (``encoder/apps/batch.cpp:740+``) runs one encoder per host
thread over a fixed file list.
"""
import threading
import time
from collections import deque
from concurrent.futures import Future

import av1gpu_amd as A


class Request:
    """One image to encode: host planes (Y, U, V numpy arrays, visible size) and its qindex."""

    def __init__(self, planes, W, H, qindex=120):
        self.planes, self.W, self.H, self.qindex = planes, W, H, qindex
        self.bytes = int(A.lib().av1mi_frame_bytes(W, H, 8, 0))
        self.t_submit = self.t_start = self.t_done = 0.0
        self.future = Future()

    @property
    def latency(self):
        return self.t_done - self.t_submit


class Server:
    def __init__(self, workers=4, hbm_budget=64 << 30, max_batch=32, qindex=120, params=None, shortlist=None):
        """params: GpParams of the search; shortlist: (network parameters, confidence) of the learned pruning, or None
        for the exhaustive search."""
        self.workers, self.budget, self.max_batch = workers, hbm_budget, max_batch
        self._pending = deque()
        self._lock = threading.Condition()
        self._inflight = 0
        self._stop = False
        self._threads = []
        self._encoders = []
        self.batches = []   # (worker, number of requests, MP) per encode, for reporting
        # each worker's persistent closed loop holds one compute unit per workgroup for the whole kernel, so the
        # workers together keep half of the device for the other workers' analysis kernels
        cus = A.device_compute_units()
        per = max(8, cus // (2 * workers)) if cus else 0
        for w in range(workers):
            s = A.C.c_void_p()
            if A.lib().av1mi_stream_create(A.C.byref(s)) != 0:
                raise RuntimeError("stream creation failed")
            cfg = A.GpConfig(qindex=qindex, queueBlocks=per)
            if params is not None:
                cfg.P = params
            enc = A.GpEncoder().init(cfg)
            if shortlist is not None:
                from av1gpu_amd import learn
                learn.set_shortlist(enc, shortlist[0], shortlist[1])
            self._encoders.append((s.value, enc))
            t = threading.Thread(target=self._run, args=(w,), daemon=True)
            self._threads.append(t)
            t.start()

    def submit(self, req: Request) -> Future:
        if req.bytes > self.budget:
            req.future.set_exception(MemoryError("image of %dx%d needs %d device bytes, over the budget %d"
                                                 % (req.W, req.H, req.bytes, self.budget)))
            return req.future
        with self._lock:
            req.t_submit = time.perf_counter()
            self._pending.append(req)
            self._lock.notify()
        return req.future

    def _take(self):
        """Admitted requests for one batch (oldest first, within the memory budget), or None when stopping."""
        with self._lock:
            while True:
                if self._stop and not self._pending:
                    return None
                batch, need = [], 0
                while self._pending and len(batch) < self.max_batch:
                    r = self._pending[0]
                    if self._inflight + need + r.bytes > self.budget:
                        break
                    batch.append(self._pending.popleft())
                    need += r.bytes
                if batch:
                    self._inflight += need
                    return batch
                self._lock.wait()

    def _run(self, w):
        stream, enc = self._encoders[w]
        while True:
            batch = self._take()
            if batch is None:
                return
            t0 = time.perf_counter()
            frames = []
            try:
                for r in batch:
                    r.t_start = t0
                    frames.append(A.DeviceFrame(r.planes, r.W, r.H, qindex=r.qindex, stream=stream))
                out, _ = enc.encode(frames, stream=stream, want_recon=True)
                for f in frames:
                    f.close()
                A.lib().av1mi_stream_sync(stream)
                t1 = time.perf_counter()
                for r, o in zip(batch, out):
                    r.t_done = t1
                    r.future.set_result(o)
                self.batches.append((w, len(batch), sum(r.W * r.H for r in batch) / 1e6))
            except Exception as e:   # the whole batch fails together
                for f in frames:
                    f.close()
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(e)
            finally:
                with self._lock:
                    self._inflight -= sum(r.bytes for r in batch)
                    self._lock.notify_all()

    def close(self):
        with self._lock:
            self._stop = True
            self._lock.notify_all()
        for t in self._threads:
            t.join()
        for s, enc in self._encoders:
            enc.close()
            A.lib().av1mi_stream_destroy(s)
        self._encoders = []


def poisson_run(server, images, rate, seed=7):
    """Submit images [(planes, W, H, qindex)] with exponential inter-arrival times of mean 1 / rate; wait for all.
    Returns the requests (with latencies) and the wall time from the first arrival to the last completion."""
    import numpy as np
    rng = np.random.default_rng(seed)
    reqs = [Request(p, W, H, q) for p, W, H, q in images]
    t0 = time.perf_counter()
    for r in reqs:
        server.submit(r)
        time.sleep(rng.exponential(1.0 / rate))
    for r in reqs:
        r.future.result()
    return reqs, max(r.t_done for r in reqs) - t0


def serving_spec(n, seed=5000):
    """Serving workload: n single images, sizes log-uniform over 0.2-24 MP with 1 % of them 24 MP, aspect ratio
    uniform over [0.5, 2], sides multiples of 8, qindex 120; image i uses synthetic seed seed + i."""
    import math
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        mp = 24.0 if rng.uniform() < 0.01 else math.exp(rng.uniform(math.log(0.2), math.log(24.0)))
        a = rng.uniform(0.5, 2.0)
        W = max(8, int(round(math.sqrt(mp * 1e6 * a) / 8.0)) * 8)
        H = max(8, int(round(math.sqrt(mp * 1e6 / a) / 8.0)) * 8)
        out.append((W, H, 120, seed + i))
    return out
