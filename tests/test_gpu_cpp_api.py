"""The reference-shaped C++ API (include/av1mi_gp.hpp: GpEncoder / GpConfig / GpFrameIn / GpFrameOut / GpTimes /
GpDevRecords, gp_check_partitions, gp_frames_to_10bit) on the GPU, through its compiled test driver
(tests/cpp/gp_encoder_check.cpp): its records equal the C ABI's, the frame parameters and the cropped reconstruction are
the reference's, device-resident records hold the same bytes, the 10-bit conversion maps every sample, and options
outside the core are refused."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "gp_encoder_check")


def test_gpu_cpp_api():
    if not os.path.exists(EXE):
        raise RuntimeError("tests/cpp/gp_encoder_check is not built (run __graft_entry__.build())")
    print("running", EXE, flush=True)
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0 and "gp ok" in out.stdout
