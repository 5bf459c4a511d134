"""The run-of-samples intra predictor (pred_col, av1gpu_amd/csrc/av1m_common.h) equals the per-sample one (pred_px)
for every mode, angle delta, block size, edge upsampling and run shape: tests/cpp/pred_col_check.cpp, built with the
host compiler."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pred_col_equals_pred_px(tmp_path):
    exe = str(tmp_path / "pred_col_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "cpp", "pred_col_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ") and int(r.stdout.split()[1]) > 100_000_000
