"""Serving loop (continuous batching over several workers and streams): every request's record and reconstruction
equal the same image encoded alone, whatever it was batched or overlapped with, also when the HBM budget admits only
one image at a time."""
import threading

import numpy as np
import pytest

import av1gpu_amd as A
from av1gpu_amd import learn, serve

pytestmark = pytest.mark.gpu


def _alone(specs, shortlist=None):
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
    if shortlist is not None:
        learn.set_shortlist(enc, *shortlist)
    out = []
    for W, H, q, s in specs:
        f = A.DeviceFrame(A.synth_image(s, W, H), W, H, qindex=q)
        out.append(enc.encode([f])[0][0])
        f.close()
    enc.close()
    return out


def _same(a, b):
    assert np.array_equal(a.blocks, b.blocks) and np.array_equal(a.txb, b.txb) and np.array_equal(a.coeffs, b.coeffs)
    for p in range(3):
        assert np.array_equal(a.recon[p], b.recon[p])


SPECS = [(200, 136, 80, 7001), (96, 64, 120, 7002), (320, 240, 160, 7003), (64, 200, 120, 7004), (136, 72, 120, 7005),
         (264, 200, 80, 7006)]


@pytest.mark.parametrize("workers,budget", [(3, 8 << 30), (2, None)])
def test_serving_results_equal_alone(workers, budget):
    ref = _alone(SPECS)
    if budget is None:   # room for the largest single request only: admission serialises
        budget = max(A.lib().av1mi_frame_bytes(W, H, 8, 0) for W, H, _, _ in SPECS)
    srv = serve.Server(workers=workers, hbm_budget=budget, max_batch=3)
    reqs = [serve.Request(A.synth_image(s, W, H), W, H, q) for W, H, q, s in SPECS]
    # submit from two threads at once
    th = [threading.Thread(target=lambda rs: [srv.submit(r) for r in rs], args=(reqs[k::2],)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    outs = [r.future.result(timeout=120) for r in reqs]
    srv.close()
    for a, b in zip(ref, outs):
        _same(a, b)
    assert all(r.latency > 0 for r in reqs)


def test_serving_with_learned_pruning_equals_alone():
    """The serving loop on the batched+learned path: each request equals its image encoded alone with the same network."""
    sl = (learn.load_params(), 0.8)
    ref = _alone(SPECS, sl)
    srv = serve.Server(workers=2, hbm_budget=8 << 30, max_batch=4, shortlist=sl)
    reqs = [serve.Request(A.synth_image(s, W, H), W, H, q) for W, H, q, s in SPECS]
    for r in reqs:
        srv.submit(r)
    outs = [r.future.result(timeout=120) for r in reqs]
    srv.close()
    for a, b in zip(ref, outs):
        _same(a, b)


def test_serving_rejects_oversized_request():
    srv = serve.Server(workers=1, hbm_budget=1 << 20)
    r = serve.Request(A.synth_image(1, 256, 256), 256, 256, 120)
    with pytest.raises(MemoryError):
        srv.submit(r).result(timeout=10)
    srv.close()
