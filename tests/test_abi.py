"""The C-ABI library loads without a GPU and exports every entry point include/av1mi.h declares; the host mirror's
structures have the C layouts.  No compute calls here."""
import ctypes as C
import os
import re

import av1gpu_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "av1mi.h")).read()
    names = sorted(set(re.findall(r"\b(av1mi_\w+)\s*\(", hdr)))
    assert len(names) >= 14
    lib = C.CDLL(A.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_record_layouts():
    assert A.BLOCK_DTYPE.itemsize == 16 and A.TXB_DTYPE.itemsize == 16
    assert C.sizeof(A._Params) == 6 * 4 + 13 * 4 + 3 * 4


def test_default_params_match_c():
    p = A._Params()
    A.lib().av1mi_default_params(C.byref(p))
    d = A.GpParams()
    for n, _ in A._Params._fields_:
        assert abs(getattr(p, n) - getattr(d, n)) < 1e-7, n


def test_partition_check_refuses_options_without_a_partition():
    A.check_partitions(A.GpParams())
    bad = A.GpParams(shapeMask=1 << 12, maxPart=32)   # 64x64 only, but nodes above 32 may not use NONE
    try:
        A.check_partitions(bad)
        assert False, "expected a refusal"
    except ValueError:
        pass
