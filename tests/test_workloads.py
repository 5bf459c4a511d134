"""The benchmark workload generators: the mixed-size batch and the serving mix follow their stated distributions and
are deterministic."""
import math

import numpy as np

import av1gpu_amd as A
from av1gpu_amd import serve


def test_batch_spec():
    s = A.batch_spec(256)
    assert s == A.batch_spec(256) and len(s) == 256
    mp = np.array([W * H / 1e6 for W, H, _, _ in s])
    assert all(W % 8 == 0 and H % 8 == 0 for W, H, _, _ in s)
    assert mp.min() > 0.17 and mp.max() < 24.5
    assert all(0.45 < W / H < 2.2 for W, H, _, _ in s)
    assert {q for _, _, q, _ in s} == {80, 120, 160}
    # log-uniform: the median near sqrt(0.18 * 24) MP
    assert abs(math.log(np.median(mp)) - 0.5 * math.log(0.18 * 24)) < 0.4
    assert [sd for _, _, _, sd in s] == list(range(1000, 1256))


def test_serving_spec():
    s = serve.serving_spec(2000)
    assert s == serve.serving_spec(2000)
    mp = np.array([W * H / 1e6 for W, H, _, _ in s])
    big = np.sum(np.abs(mp - 24.0) < 0.3)
    assert 5 <= big <= 45          # about 1 % of requests are 24 MP images
    assert mp.min() > 0.18 and mp.max() < 24.5 and all(q == 120 for _, _, q, _ in s)
