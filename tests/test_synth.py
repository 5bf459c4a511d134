"""The synthetic generator reproduces the committed plane hashes (tests/golden/synth_sha256.json) and its statistics
(luma pink noise normalised to std 40 before the edges, chroma std 12 around 128)."""
import json
import os

import numpy as np

import av1gpu_amd as A
from golden.make_synth_golden import digest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_synth_matches_golden_hashes():
    for c in json.load(open(os.path.join(HERE, "golden", "synth_sha256.json"))):
        assert digest(c["seed"], c["W"], c["H"], c["ssx"], c["bd"]) == c["sha256"], c["seed"]


def test_synth_statistics():
    y, u, v = A.synth_image(1000, 512, 512)
    assert 100 < y.mean() < 160 and y.std() > 35
    for c in (u, v):
        assert abs(c.mean() - 128) < 4 and 9 < c.std() < 15
    y10 = A.synth_image(1000, 64, 64, bd=10)[0]
    assert y10.dtype == np.uint16 and y10.max() <= 1023
