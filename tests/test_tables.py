"""Spec tables and derived tables: the generated AV1 tables against the reference's headers and the mathematics, and the
product's host-built device tables against the oracle's, bit for bit."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import av1gpu_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OURS = os.path.join(ROOT, "av1gpu_amd", "csrc", "av1_spec_tables.h")
REF_INC = "/root/reference/av1core/include/av1g"


def _ours():
    s = open(OURS).read()
    return {n: [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
            for _t, n, _d, body in re.findall(r"inline constexpr (\w+) (\w+)((?:\[\d+\])+) = \{([^}]*)\};", s)}


def test_spec_tables_match_reference_headers():
    if not os.path.isdir(REF_INC):
        pytest.skip("reference tree not present")
    ref = {}
    for f in ("spec_tables.h", "cdf_defaults.h"):
        s = open(os.path.join(REF_INC, f)).read()
        for _t, n, _d, body in re.findall(r"AV1G_TABLE\((\w+), (\w+), ((?:\[\d+\])+), \{([^}]*)\}\)", s):
            ref[n] = [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    ours = _ours()
    assert len(ours) > 90
    for n, v in ours.items():
        assert v == ref[n], n


def test_spec_tables_closed_forms():
    t = _ours()
    # Cos128_Lookup[a] = round(4096 cos(a pi / 128))
    assert t["Cos128_Lookup"] == [round(4096 * math.cos(a * math.pi / 128)) for a in range(65)]
    # every scan is a permutation of its coefficient positions
    for n, v in t.items():
        if "_Scan_" in n:
            assert sorted(v) == list(range(len(v))), n
    # default CDFs: non-decreasing, every CDF ends at 32768 followed by the counter 0
    for n, v in t.items():
        if n.endswith("_Cdf"):
            assert all(x <= 32768 for x in v), n
    # quantizer lookups: monotone, the spec's end points
    dc, ac = t["Dc_Qlookup"], t["Ac_Qlookup"]
    assert dc[0] == 4 and ac[0] == 4 and dc[255] == 1336 and ac[255] == 1828
    for q in (dc, ac):
        assert all(q[i] <= q[i + 1] for i in range(255))
    # transform sizes: widths / heights are 4 << log2
    assert all(w == 1 << l for w, l in zip(t["Tx_Width"], t["Tx_Width_Log2"]))
    assert all(h == 1 << l for h, l in zip(t["Tx_Height"], t["Tx_Height_Log2"]))


def _misc(f):
    arrs = [np.zeros(n, np.float32) for n in (19 * 16, 19, 13, 50, 1, 512, 512)]
    f(*[a.ctypes.data for a in arrs])
    return arrs


def test_device_tables_equal_oracle(oracle):
    L = A.lib()
    for k in range(3):
        for ln in range(2, 7):
            a, b = np.zeros(4096, np.float32), np.zeros(4096, np.float32)
            na, nb = L.av1mi_fwd_matrix(k, ln, a.ctypes.data), oracle.lib.orc_fwd_matrix(k, ln, b.ctypes.data)
            assert na == nb
            assert np.array_equal(a[:na].view(np.uint32), b[:nb].view(np.uint32)), (k, ln)
    for x, y in zip(_misc(L.av1mi_misc_tables), _misc(oracle.lib.orc_misc_tables)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    for q in range(4):
        a, b = np.zeros(20000, np.float32), np.zeros(20000, np.float32)
        na, nb = L.av1mi_cost_table(q, a.ctypes.data), oracle.lib.orc_cost_table(q, b.ctypes.data)
        assert na == nb and np.array_equal(a[:na].view(np.uint32), b[:nb].view(np.uint32))


def test_forward_matrices_invert_the_exact_inverse(oracle):
    """F . M = I: the forward matrices invert the impulse responses of the exact 1-D inverse transforms."""
    oracle.lib.orc_inv_1d.restype = C.c_uint32
    oracle.lib.orc_inv_1d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    for kind, lns in ((0, range(2, 7)), (1, range(2, 5)), (2, range(2, 6))):
        for ln in lns:
            n = 1 << ln
            M = np.zeros((n, n))
            for j in range(n):
                t = np.zeros(64, np.int32)
                t[j] = 1 << 14
                oracle.lib.orc_inv_1d(t.ctypes.data, kind, ln, 31)
                M[:, j] = t[:n] / 16384.0
            F = np.zeros(n * n, np.float32)
            assert oracle.lib.orc_fwd_matrix(kind, ln, F.ctypes.data) == n * n
            err = np.abs(F.reshape(n, n).astype(np.float64) @ M - np.eye(n)).max()
            assert err < 2e-6, (kind, ln, err)


def test_lambda_closed_form(oracle):
    lam = _misc(oracle.lib.orc_misc_tables)[5].reshape(2, 256)
    tabs = _ours()
    for q in (1, 60, 120, 160, 255):
        dq = tabs["Dc_Qlookup"][q]
        w = min(max(((255 - q) * 3) // 4, 0), 72)
        expect = dq * dq * (3.3 + 0.0015 * dq) * (128 + w) / 128.0 / 2048.0
        assert abs(lam[0, q] - expect) <= 1e-6 * expect
