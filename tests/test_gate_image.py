"""The reference's decoder-gate picture (tests/gate_images/cfl_edge_211x149.png: odd sides, CfL blocks with a split
luma tx on the bottom frame edge), converted once to YUV 4:2:0 by tests/golden/make_cfl_edge.py: the oracle's record of
it, at the gate's qindex 60 and 120 with every partition and with AB-only partitions at the larger nodes (the gate's
--partN 0xF0, CMakeLists.txt:244-252), is decoded by the reference's backend and by dav1d to the oracle's
reconstruction.  The GPU side of the same cases is in test_gpu_fullsize.py."""
import os

import numpy as np
import pytest

import refcheck

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfl_edge_211x149_yuv420.npz")
W, H = 211, 149


def gate_planes():
    d = np.load(FIXTURE)
    return [d["y"], d["u"], d["v"]]


def test_fixture_shapes():
    y, u, v = gate_planes()
    assert y.shape == (H, W) and u.shape == ((H + 1) // 2, (W + 1) // 2) and v.shape == u.shape
    assert y.dtype == np.uint8 and u.dtype == np.uint8


@pytest.mark.parametrize("q", [60, 120])
@pytest.mark.parametrize("part_n", [0x3FF, 0xF0])
def test_oracle_gate_image_decodes(oracle, ref, q, part_n):
    planes = gate_planes()
    r = oracle.encode(planes, W, H, qindex=q, params=oracle.params(partMaskN=part_n), threads=4)
    b = r["blocks"]
    assert (b["uv_mode"] == 13).any()                      # CfL blocks are present
    assert (b["mi_row"] * 4 + 4 * 2 > H - 8).any()          # blocks on the bottom frame edge
    dec, avif, _ = ref.encode(r, W, H, q, avif=True)
    for p in range(3):
        assert np.array_equal(r["recon"][p][:dec[p].shape[0], :dec[p].shape[1]], dec[p]), p
    if not refcheck.dav1d_available():
        pytest.skip("dav1d (Pillow's libavif) is not available: the reference backend's decode was checked")
    d = refcheck.dav1d_decode_yuv(avif)
    for p in range(3):
        assert np.array_equal(d[p], dec[p]), p
