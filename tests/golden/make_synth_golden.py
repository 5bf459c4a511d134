"""Writes synth_sha256.json: SHA-256 of the synthetic generator's planes for seeds 1000..1003 (the benchmark
inputs), so any machine can check it regenerates the same images.  Run from the repo root after building."""
import hashlib
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
import av1gpu_amd as A  # noqa: E402

CASES = [(1000, 320, 240, 1, 8), (1001, 257, 131, 1, 8), (1002, 512, 512, 1, 8), (1003, 200, 136, 1, 10)]


def digest(seed, W, H, ssx, bd):
    return [hashlib.sha256(p.tobytes()).hexdigest() for p in A.synth_image(seed, W, H, ssx, bd)]


if __name__ == "__main__":
    out = [dict(seed=s, W=W, H=H, ssx=x, bd=b, sha256=digest(s, W, H, x, b)) for s, W, H, x, b in CASES]
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "synth_sha256.json"), "w") as f:
        json.dump(out, f, indent=1)
