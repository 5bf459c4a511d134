"""Training labels of the pruning network (av1gpu_amd/learn.py sb_labels / sb_tiles) against a raster restatement
over oracle records: every 4x4 unit of the coded area is painted with the size and luma mode of the block that covers
it (which also checks that the blocks tile the area exactly once, pinning the block-size tables), and a node's expand
label is then "some unit inside it belongs to a block at most half the node's size", its own label "some unit inside it
belongs to a block whose larger side is the node's size and at most twice its smaller side", its four label the same
with the larger side 4 times the smaller (HORZ_4 / VERT_4 shapes), its mode labels the modes that appear among the
painted units."""
import numpy as np
import pytest

import av1gpu_amd as A
from av1gpu_amd import learn as Ln

W4 = (1, 1, 2, 2, 2, 4, 4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 1, 4, 2, 8, 4, 16)   # block width in 4x4 units per bsize
H4 = (1, 2, 1, 2, 4, 2, 4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 4, 1, 8, 2, 16, 4)


def _paint(blocks, mic, mir):
    size = np.zeros((mir, mic), np.int64)
    ratio = np.zeros((mir, mic), np.int64)
    mode = np.full((mir, mic), -1, np.int64)
    cover = np.zeros((mir, mic), np.int64)
    for b in blocks:
        r, c, s = int(b["mi_row"]), int(b["mi_col"]), int(b["bsize"])
        h, w = min(H4[s], mir - r), min(W4[s], mic - c)
        size[r:r + h, c:c + w] = 4 * max(W4[s], H4[s])
        ratio[r:r + h, c:c + w] = max(W4[s], H4[s]) // min(W4[s], H4[s])
        mode[r:r + h, c:c + w] = int(b["y_mode"])
        cover[r:r + h, c:c + w] += 1
    return size, ratio, mode, cover


@pytest.mark.parametrize("W,H,q,seed", [(200, 136, 120, 41), (264, 200, 40, 42), (72, 264, 200, 43)])
def test_labels_match_raster_restatement(oracle, W, H, q, seed):
    planes = A.synth_image(seed, W, H)
    r = oracle.encode(A.pad_planes(planes, W, H)[0], W, H, qindex=q, threads=8)
    mic, mir = r["MiCols"], r["MiRows"]
    size, ratio, mode, cover = _paint(r["blocks"], mic, mir)
    assert (cover == 1).all()                      # the record's blocks tile the coded area exactly once
    lab, msk = Ln.sb_labels(r["blocks"], W, H)
    sbc, sbr = (mic + 15) // 16, (mir + 15) // 16
    assert lab.shape == (sbr * sbc, Ln.N_OUT) and msk.shape == lab.shape
    for s in range(sbr * sbc):
        r0, c0 = (s // sbc) * 16, (s % sbc) * 16
        for lv in range(4):
            n4 = 16 >> lv
            for ny in range(1 << lv):
                for nx in range(1 << lv):
                    node = Ln.LEVEL_BASE[lv] + ny * (1 << lv) + nx
                    y0, x0 = r0 + ny * n4, c0 + nx * n4
                    inside = y0 + n4 <= mir and x0 + n4 <= mic
                    assert msk[s, node] == (1.0 if inside else 0.0)
                    assert msk[s, Ln.OWN0 + node] == msk[s, node] and msk[s, Ln.FOUR0 + node] == msk[s, node]
                    sub = size[y0:y0 + n4, x0:x0 + n4]
                    rat = ratio[y0:y0 + n4, x0:x0 + n4]
                    if sub.size:
                        assert lab[s, node] == (1.0 if sub.min() * 2 <= 4 * n4 else 0.0), (s, lv, ny, nx)
                        own = ((sub == 4 * n4) & (rat <= 2)).any()
                        four = ((sub == 4 * n4) & (rat == 4)).any()
                        assert lab[s, Ln.OWN0 + node] == (1.0 if own else 0.0), (s, lv, ny, nx)
                        assert lab[s, Ln.FOUR0 + node] == (1.0 if four else 0.0), (s, lv, ny, nx)
                        if lv == 3:
                            assert not four              # 8x8 nodes have no 4-way partitions
                        if sub.max() <= 4 * n4:        # not inside a larger block: children, own or 4-way shapes used
                            assert lab[s, node] + lab[s, Ln.OWN0 + node] + lab[s, Ln.FOUR0 + node] >= 1.0
                    if lv and lab[s, node]:            # children used implies the parent's children used
                        parent = Ln.LEVEL_BASE[lv - 1] + (ny // 2) * (1 << (lv - 1)) + nx // 2
                        assert lab[s, parent] == 1.0
        m = mode[r0:r0 + 16, c0:c0 + 16].ravel()
        used = (np.bincount(m, minlength=13)[:13] > 0).astype(np.float32)
        assert np.array_equal(lab[s, Ln.MODE0:], used)


def test_tiles_are_the_padded_superblocks():
    W, H = 136, 72
    y = A.synth_image(44, W, H)[0]
    t = Ln.sb_tiles(y, W, H)
    pad = A.pad_planes([y, y[:36, :68], y[:36, :68]], W, H)[0][0]    # edge replication as the encoder's planes
    assert t.shape == (3 * 2, 4096)
    for s in range(6):
        r, c = s // 3, s % 3
        assert np.array_equal(t[s].reshape(64, 64), pad[64 * r:64 * r + 64, 64 * c:64 * c + 64])


def test_bd_rate_closed_forms():
    """A curve against itself is 0 %; a curve with every rate scaled by k (same PSNR) is (k - 1) * 100 %."""
    r = np.array([1e5, 2e5, 4e5, 8e5])
    p = np.array([30.0, 33.5, 36.8, 39.9])
    assert abs(Ln.bd_rate(r, p, r, p)) < 1e-9
    assert abs(Ln.bd_rate(r, p, 1.1 * r, p) - 10.0) < 1e-6
    assert abs(Ln.bd_rate(r, p, 0.8 * r, p) + 20.0) < 1e-6
