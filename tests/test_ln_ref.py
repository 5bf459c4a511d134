"""The float64 restatement of the pruning network (tests/ln_ref.py) pinned by closed forms: BF16 rounding on known
values, a one-tap convolution equal to the stride-2 sampling of its input, and the backward restatement equal to central
finite differences of the forward restatement (BF16 rounding switched off) for sampled parameters of every layer."""
import numpy as np
import pytest

import ln_ref as R
from av1gpu_amd import learn as Ln


def test_bf16_rounding_known_values():
    # ties to even both ways: 1 + 2^-8 -> 1 (even), 1 + 3 * 2^-8 -> 1 + 2^-6 (odd last bit rounds up)
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 1 + 2 ** -7, 65504.0, 1 + 2 ** -9], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, -2.5, 1 + 2 ** -7, 65536.0, 1.0], np.float32)
    assert np.array_equal(R.bf16(x), want)


def test_one_tap_convolution_samples_the_input():
    rng = np.random.default_rng(1)
    x = R.bf16(rng.uniform(-1, 1, (2, 16, 16, 32)).astype(np.float32))
    w = np.zeros((32, Ln.KP[1]), np.float32)
    b = np.zeros(32, np.float32)
    for co in range(32):
        w[co, 4 * 32 + (co + 5) % 32] = 1.0                        # centre tap (ky = kx = 1), input channel co + 5
    y = R.ref_conv(x, w, b, 1)
    want = np.maximum(x[:, ::2, ::2, :][..., (np.arange(32) + 5) % 32], 0)
    assert np.array_equal(y, want)


@pytest.fixture
def exact(monkeypatch):
    monkeypatch.setattr(R, "bf16", lambda a: np.asarray(a, np.float64))   # a float64 forward, no rounding


def _forward(p, tiles, q):
    convs, w5, b5 = Ln.unpack(p)
    acts = [R.input_tensor(tiles, q)]
    for l in range(4):
        acts.append(R.ref_conv(acts[-1], convs[l][0], convs[l][1], l))
    return acts


def test_backward_equals_finite_differences(exact):
    rng = np.random.default_rng(2)
    p = Ln.init_params(3).astype(np.float64)
    tiles = rng.integers(0, 256, (2, 4096)).astype(np.uint8)
    q = np.array([80, 160], np.uint8)
    lab = (rng.uniform(size=(2, Ln.N_OUT)) < 0.4).astype(np.float64)
    msk = (rng.uniform(size=(2, Ln.N_OUT)) < 0.9).astype(np.float64)
    acts = _forward(p, tiles, q)
    loss, g = R.ref_grads(p, acts, lab, msk)
    convs, w5, b5 = Ln.unpack(np.arange(len(p)))                    # index views: where each parameter lives
    picks = [int(w5[3, 7]), int(w5[90, 300]), int(b5[100]), int(convs[3][0][5, 40]), int(convs[2][1][9]),
             int(convs[1][0][17, 200]), int(convs[0][0][2, 11])]
    for i in picks:
        eps = 1e-6 * max(1.0, abs(p[i]))
        lo, hi = p.copy(), p.copy()
        lo[i] -= eps
        hi[i] += eps
        l_lo = R.ref_grads(lo, _forward(lo, tiles, q), lab, msk)[0]
        l_hi = R.ref_grads(hi, _forward(hi, tiles, q), lab, msk)[0]
        fd = (l_hi - l_lo) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-4 * max(abs(fd), abs(g[i])) + 1e-8, (i, fd, g[i])
