"""The reference-side binding compiles against the reference's own headers (encoder/src/gp.h, av1core/include/av1g/
decisions.h) with g++ and no CUDA: this library's record entries and frame parameters have the reference's layouts,
its C++ GpParams the reference's fields in the same order, every mapped field is carried by name and every knob the core
lacks is refused (tests/cpp/binding_check.cpp).  Skipped where the reference tree is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


def test_binding_compiles_against_reference_headers(tmp_path):
    if not os.path.isdir(os.path.join(REF, "encoder", "src")):
        pytest.skip("reference tree not present")
    exe = str(tmp_path / "binding_check")
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                           "-I" + os.path.join(REF, "encoder", "src"), "-I" + os.path.join(REF, "av1core", "include"),
                           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "binding_check.cpp"),
                           "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "binding ok" in out.stdout, out.stdout + out.stderr
