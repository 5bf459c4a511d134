"""The oracle's whole search on seeded synthetic images, checked by the reference's own CPU backend (oracle/_ref:
av1g_layout, av1g_validate, av1g_encode) -- the record is legal, its tx layout is the decoder's, and the oracle's
reconstruction equals the decoder-exact reconstruction of its record -- and by dav1d (Pillow's libavif) decoding the
AVIF the reference writer makes from the record."""
import numpy as np
import pytest

import av1gpu_amd as A
import refcheck


@pytest.mark.parametrize("W,H,seed,q", [(64, 64, 1, 120), (96, 80, 2, 80), (130, 66, 3, 160), (8, 8, 4, 120), (200, 136, 5, 255)])
def test_oracle_record_decodes_to_its_reconstruction(oracle, ref, W, H, seed, q):
    planes = A.synth_image(1000 + seed, W, H)
    r = oracle.encode(planes, W, H, qindex=q)
    dec, avif, info = ref.encode(r, W, H, q, avif=True)
    for p in range(3):
        mine = r["recon"][p][:dec[p].shape[0], :dec[p].shape[1]]
        assert np.array_equal(mine, dec[p]), p
    assert np.array_equal(info["eob"], r["txb"]["eob"])
    assert info["n_fixed"] == 0
    if refcheck.dav1d_available():
        d = refcheck.dav1d_decode_yuv(avif)
        for p in range(3):
            assert np.array_equal(d[p], dec[p]), p


def test_oracle_10bit_record_decodes(oracle, ref):
    W, H = 72, 56
    planes = A.synth_image(1010, W, H, bd=10)
    r = oracle.encode(planes, W, H, bd=10, qindex=120)
    dec, avif, _ = ref.encode(r, W, H, 120, bd=10, avif=True)
    for p in range(3):
        assert np.array_equal(r["recon"][p][:dec[p].shape[0], :dec[p].shape[1]], dec[p])


@pytest.mark.parametrize("W,H,seed,q,bd", [(200, 136, 1, 120, 8), (256, 256, 11, 255, 8), (200, 264, 13, 240, 10), (130, 18, 4, 200, 8)])
def test_oracle_444_record_decodes(oracle, ref, W, H, seed, q, bd):
    """4:4:4: chroma at full resolution, blocks with a 64-sample side carrying 2 or 4
    chroma tx blocks (the high-qindex cases hold such blocks); the reference's backend reconstructs the record to the
    oracle's reconstruction and dav1d decodes the reference writer's AVIF of it to the same."""
    planes = A.synth_image(seed, W, H, ssx=0, bd=bd)
    r = oracle.encode(planes, W, H, ssx=0, bd=bd, qindex=q, threads=4)
    if q >= 240:
        assert np.isin(r["blocks"]["bsize"], [10, 11, 12, 20, 21]).any()
    dec, avif, info = ref.encode(r, W, H, q, ssx=0, bd=bd, avif=True)
    for p in range(3):
        assert np.array_equal(r["recon"][p][:dec[p].shape[0], :dec[p].shape[1]], dec[p]), p
    assert np.array_equal(info["eob"], r["txb"]["eob"])
    if refcheck.dav1d_available():
        d = refcheck.dav1d_decode_yuv(avif)
        for p in range(3):
            assert np.array_equal(d[p], dec[p]), p


def test_oracle_threads_do_not_change_the_result(oracle):
    W, H = 160, 128
    planes = A.synth_image(1020, W, H)
    a = oracle.encode(planes, W, H, threads=1)
    b = oracle.encode(planes, W, H, threads=4)
    for k in ("blocks", "block_txb", "txb", "coeffs"):
        assert np.array_equal(a[k], b[k]), k


def test_oracle_partition_leaves_tile_every_superblock(oracle):
    """The DP's leaves cover each SB's in-frame area exactly once (and the record's blocks, in decode order)."""
    W, H = 136, 72
    r = oracle.encode(A.synth_image(1030, W, H), W, H, detail=True)
    bw = [1, 1, 2, 2, 2, 4, 4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 1, 4, 2, 8, 4, 16]
    bh = [1, 2, 1, 2, 4, 2, 4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 4, 1, 8, 2, 16, 4]
    mic, mir = r["MiCols"], r["MiRows"]
    cover = np.zeros((mir, mic), np.int32)
    for b in r["blocks"]:
        cover[b["mi_row"]:b["mi_row"] + bh[b["bsize"]], b["mi_col"]:b["mi_col"] + bw[b["bsize"]]] += 1
    assert (cover == 1).all()


def test_oracle_dp_matches_brute_force_on_8x8_nodes(oracle):
    """At an 8x8 node the DP's choice is the cheapest of NONE / HORZ / VERT / SPLIT by brute force over the analysis
    records (J + lambda * averaged partition bits)."""
    W, H = 64, 64
    r = oracle.encode(A.synth_image(1040, W, H), W, H, detail=True)
    br = r["blockres"]
    leaves = r["leaves"][0][:r["leaf_count"][0]]
    # every 8x8 area covered by leaves smaller than 8x8 must be a legal 8x8 partition: 1, 2 or 4 leaves
    small = [l for l in leaves if l["bsize"] in (0, 1, 2)]
    areas = {}
    for l in small:
        areas.setdefault((l["x4"] // 2, l["y4"] // 2), []).append(int(l["bsize"]))
    for k, v in areas.items():
        assert sorted(v) in ([0, 0, 0, 0], [1, 1], [2, 2]), (k, v)
    assert all(np.isfinite(br[s]["J"][0]) for s in br if len(br[s]))
