"""The oracle's analysis hooks (the reference's external screener GpBatch::extCand, encoder/src/gp.h:136-143 and
164-171, read at encoder/src/k_analysis.cuh:586-608; and its block lists GpAnX, encoder/src/an_stage.h:23-27), pinned by
equivalences that hold by construction of the search:
* external candidates equal to the exhaustive analysis's own candidates (winner first) reproduce its records;
* an external list with no usable entry (modes of 13 or more) falls back to the screening, so records do not change;
* a single usable entry leaves that block exactly that candidate;
* a block list naming every block changes nothing, and an empty list for a shape equals taking the shape out of
  shapeMask."""
import numpy as np
import pytest

import av1gpu_amd as A
from refcheck import EXT_DTYPE

W, H, SEED, Q = 136, 72, 50, 120
SHAPES = [s for s in range(22) if not 13 <= s <= 15]
BW = {0: (4, 4), 1: (4, 8), 2: (8, 4), 3: (8, 8), 4: (8, 16), 5: (16, 8), 6: (16, 16), 7: (16, 32), 8: (32, 16), 9: (32, 32),
      10: (32, 64), 11: (64, 32), 12: (64, 64), 16: (4, 16), 17: (16, 4), 18: (8, 32), 19: (32, 8), 20: (16, 64), 21: (64, 16)}


def grid(s, w, h):
    mic, mir = 2 * ((w + 7) >> 3), 2 * ((h + 7) >> 3)
    bw, bh = BW[s]
    return -(-mic * 4 // bw), -(-mir * 4 // bh)


def same_record(a, b):
    for k in ("blocks", "block_txb", "txb", "coeffs"):
        assert np.array_equal(a[k], b[k]), k
    for p in range(3):
        assert np.array_equal(a["recon"][p], b["recon"][p]), p


@pytest.fixture(scope="module")
def planes():
    return A.synth_image(SEED, W, H)


@pytest.fixture(scope="module")
def plain(oracle, planes):
    return oracle.encode(planes, W, H, qindex=Q, threads=4, detail=True)


def own_candidates(br, topk=6):
    e = np.zeros((len(br), topk), EXT_DTYPE)
    e["mode"] = 255
    for i, r in enumerate(br):
        for k in range(r["ncand"]):
            c = int(r["cand"][k])
            e[i, k]["mode"], e[i, k]["delta"] = c >> 3, (c & 7) - 3
    return e


def test_ext_own_candidates_reproduce_the_search(oracle, planes, plain):
    ext = {s: own_candidates(plain["blockres"][s]) for s in SHAPES}
    r = oracle.encode(planes, W, H, qindex=Q, threads=4, detail=True, ext=ext, ext_topk=6)
    same_record(plain, r)
    for s in SHAPES:
        assert np.array_equal(plain["blockres"][s], r["blockres"][s]), s


def test_ext_without_usable_entries_falls_back_to_screening(oracle, planes, plain):
    rng = np.random.default_rng(1)
    ext = {}
    for s in SHAPES:
        nx, ny = grid(s, W, H)
        e = np.zeros((nx * ny, 3), EXT_DTYPE)
        e["mode"] = rng.integers(13, 256, size=e.shape)
        ext[s] = e
    r = oracle.encode(planes, W, H, qindex=Q, threads=4, ext=ext, ext_topk=3)
    same_record(plain, r)


def test_ext_single_entry_is_the_only_candidate(oracle, planes):
    s = 3   # 8x8
    nx, ny = grid(s, W, H)
    e = np.zeros((nx * ny, 4), EXT_DTYPE)
    e["mode"] = 200
    e[:, 2]["mode"], e[:, 2]["delta"] = 3, -2           # D45 with delta -2 (directional: kept)
    e[1, 2]["mode"], e[1, 2]["delta"] = 0, 2            # DC: delta ignored
    e[2, 2]["mode"], e[2, 2]["delta"] = 1, 9            # V with an out-of-range delta: 0
    r = oracle.encode(planes, W, H, qindex=Q, threads=4, detail=True, ext={s: e}, ext_topk=4)
    br = r["blockres"][s]
    assert np.all(br["ncand"] == 1)
    assert br[0]["cand"][0] == (3 << 3 | 1) and br[1]["cand"][0] == (0 << 3 | 3) and br[2]["cand"][0] == (1 << 3 | 3)
    assert np.all(np.isfinite(br["J"]))


def test_block_list_of_every_block_changes_nothing(oracle, planes, plain):
    an = {s: np.ones(np.prod(grid(s, W, H)), np.uint8) for s in SHAPES}
    r = oracle.encode(planes, W, H, qindex=Q, threads=4, analyse=an)
    same_record(plain, r)


def test_empty_block_list_equals_shape_off(oracle, planes):
    P = oracle.params()
    P.shapeMask = 0x3F1FFF & ~((1 << 16) | (1 << 17) | (1 << 5))
    off = oracle.encode(planes, W, H, qindex=Q, threads=4, params=P, detail=True)
    an = {s: np.zeros(np.prod(grid(s, W, H)), np.uint8) for s in (16, 17, 5)}
    r = oracle.encode(planes, W, H, qindex=Q, threads=4, analyse=an, detail=True)
    same_record(off, r)
    assert np.all(r["blockres"][17]["J"] > 1e38)
