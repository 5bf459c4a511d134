"""The identities the kernels' lane-distributed forms rest on, restated lane by lane in NumPy (float32 arithmetic, one
rounding per addition as on the GPU):

* the half-group evaluation (av1gpu_amd/csrc/dev_eval.h, half_qrate / hsum2): lane l of a 16-lane half holds the
  partial sums of virtual lanes l and l + 16 of the 32-lane form and adds them first; the result equals the 32-lane
  butterfly gsum (partners lane xor 16, 8, 4, 2, 1) to the bit, for any values;
* the per-pixel SATD screening (hadamard4_lanes): 16 lanes holding a 4x4 block in raster order, butterflies with
  partners lane xor 1, 2 (rows) then 4, 8 (columns), give |coefficients| whose sum is hadamard4's result.
A wrong pairing (xor 8 before xor 16) is shown to change float sums, so the first identity is not vacuous."""
import numpy as np


def butterfly(v, order):
    """All-lanes butterfly sum: for each partner distance o in order, lane l adds the value of lane l ^ o."""
    v = np.asarray(v, np.float32).copy()
    idx = np.arange(len(v))
    for o in order:
        v = (v + v[idx ^ o]).astype(np.float32)
    return v


def gsum(v32):
    return butterfly(v32, (16, 8, 4, 2, 1))


def hsum2(a16, b16):
    v = (np.asarray(a16, np.float32) + np.asarray(b16, np.float32)).astype(np.float32)
    return butterfly(v, (8, 4, 2, 1))


def test_half_group_sum_equals_group_sum():
    rng = np.random.default_rng(5)
    for trial in range(2000):
        scale = 10.0 ** rng.integers(-3, 6)
        v = (rng.standard_normal(32) * scale).astype(np.float32)
        if trial % 3 == 0:
            v[16:] = 0.0                      # blocks of up to 16 coefficients: virtual lanes 16..31 hold 0
        g = gsum(v)
        assert (g == g[0]).all()              # every lane ends with the same value
        h = hsum2(v[:16], v[16:])
        assert (h == g[0]).all(), trial


def test_wrong_pairing_changes_sums():
    """Adding the xor-8 partner before the xor-16 one is a different tree: some float sums differ."""
    rng = np.random.default_rng(6)
    diff = 0
    for _ in range(500):
        v = (rng.standard_normal(32) * 1e3).astype(np.float32)
        diff += int(gsum(v)[0] != butterfly(v, (8, 16, 4, 2, 1))[0])
    assert diff > 0


def hadamard4(d):
    """The per-lane form: rows, then columns, then the sum of |coefficients| (dev_eval.h hadamard4)."""
    d = np.array(d, np.int64).reshape(4, 4)
    r = np.empty_like(d)
    for i in range(4):
        s0, s1, s2, s3 = d[i, 0] + d[i, 1], d[i, 0] - d[i, 1], d[i, 2] + d[i, 3], d[i, 2] - d[i, 3]
        r[i] = (s0 + s2, s1 + s3, s0 - s2, s1 - s3)
    total = 0
    for j in range(4):
        s0, s1, s2, s3 = r[0, j] + r[1, j], r[0, j] - r[1, j], r[2, j] + r[3, j], r[2, j] - r[3, j]
        total += abs(s0 + s2) + abs(s1 + s3) + abs(s0 - s2) + abs(s1 - s3)
    return total, r


def hadamard4_lanes(d):
    """The 16-lane form: lane l holds d[l] (raster); stage o: t = value of lane l ^ o, v = t - v if l & o else v + t."""
    v = np.array(d, np.int64)
    lanes = np.arange(16)
    for o in (1, 2, 4, 8):
        t = v[lanes ^ o]
        v = np.where(lanes & o, t - v, v + t)
    return np.abs(v)


def test_lane_hadamard_equals_hadamard4():
    rng = np.random.default_rng(7)
    for _ in range(3000):
        d = rng.integers(-1023, 1024, 16)
        total, _ = hadamard4(d)
        lanes = hadamard4_lanes(d)
        assert lanes.sum() == total
    # closed forms: a constant block has one nonzero coefficient 16 c; a single impulse spreads |c| over all 16
    assert hadamard4_lanes([3] * 16).sum() == 48 and (hadamard4_lanes([3] * 16) != 0).sum() == 1
    imp = [0] * 16
    imp[5] = -7
    assert (hadamard4_lanes(imp) == 7).all()
