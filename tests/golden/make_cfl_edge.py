"""Writes tests/golden/cfl_edge_211x149_yuv420.npz from the reference's decoder-gate image
tests/gate_images/cfl_edge_211x149.png (an odd-sized RGB picture whose chroma gets CfL blocks with a split luma tx on
the bottom frame edge; the reference's gate runs it at qindex 60 and 120 with AB-only partitions at the larger nodes,
CMakeLists.txt:244-252).

The conversion is the reference's 8-bit PNG path (encoder/src/imageio.cpp:339-372): BT.601 full-range YCbCr with each
sample rounded half to even and clamped to 0..255, chroma 4:2:0 as the rounded mean of each 2x2 (edge samples
repeated at odd sides).  Only the visible planes are stored (Y 211x149, U/V 106x75); the tests pad them.

Run once in the container that holds the reference tree:  python tests/golden/make_cfl_edge.py
"""
import hashlib
import os

import numpy as np
from PIL import Image

SRC = "/root/reference/tests/gate_images/cfl_edge_211x149.png"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfl_edge_211x149_yuv420.npz")


def to_yuv420(rgb):
    r, g, b = (rgb[..., k].astype(np.float64) for k in range(3))
    cl = lambda v: np.clip(np.rint(v), 0, 255).astype(np.int32)
    y = cl(0.299 * r + 0.587 * g + 0.114 * b)
    cb = cl(-0.168735892 * r - 0.331264108 * g + 0.5 * b + 128)
    cr = cl(0.5 * r - 0.418687589 * g - 0.081312411 * b + 128)
    H, W = y.shape
    ch, cw = (H + 1) // 2, (W + 1) // 2
    iy = np.minimum(np.arange(2 * ch), H - 1)
    ix = np.minimum(np.arange(2 * cw), W - 1)

    def down(c):
        e = c[iy][:, ix]
        s = e[0::2, 0::2] + e[0::2, 1::2] + e[1::2, 0::2] + e[1::2, 1::2]
        return ((s + 2) >> 2).astype(np.uint8)

    return y.astype(np.uint8), down(cb), down(cr)


def main():
    data = open(SRC, "rb").read()
    rgb = np.asarray(Image.open(SRC).convert("RGB"))
    y, u, v = to_yuv420(rgb)
    np.savez(OUT, y=y, u=u, v=v, png_sha256=np.array(hashlib.sha256(data).hexdigest()))
    print("wrote", OUT, y.shape, u.shape)


if __name__ == "__main__":
    main()
