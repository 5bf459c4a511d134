"""Host bitstream writer and AVIF muxer (av1mi_write_obus / av1mi_write_avif): the stream of a record parses back to
the same record in the reference's own OBU parser, and the AVIF decodes in dav1d (libavif) to the record's
decoder-exact reconstruction.  Records come from the oracle (the same records the GPU path emits, bit for bit)."""
import ctypes as C
import struct

import numpy as np
import pytest

import av1gpu_amd as A
import refcheck

CASES = [(64, 64, 120, 1, 8, 1), (128, 96, 80, 2, 8, 1), (200, 136, 160, 3, 8, 1), (130, 18, 30, 8, 8, 1),
         (72, 40, 255, 5, 8, 1), (136, 72, 120, 50, 10, 1), (96, 64, 120, 60, 8, 0), (4160, 64, 120, 9, 8, 1)]
# 4:4:4 (seq_profile 1): ssx = 0; the qindex-255 cases hold 64-sided blocks with several chroma tx blocks
CASES444 = [(200, 136, 120, 11, 8, 1), (256, 256, 255, 12, 8, 1), (200, 264, 240, 13, 10, 1), (130, 18, 200, 14, 8, 0)]


def _record(oracle, W, H, q, seed, bd, edge, ssx=1):
    planes = A.synth_image(1000 + seed, W, H, ssx=ssx, bd=bd)
    r = oracle.encode(planes, W, H, ssx=ssx, qindex=q, bd=bd, edge_filter=edge, threads=8)
    cols, rows = C.c_int(), C.c_int()
    cs, rs = (C.c_int * 65)(), (C.c_int * 65)()
    oracle.lib.orc_tile_layout(W, H, C.byref(cols), C.byref(rows), cs, rs)
    rec = dict(width=W, height=H, base_q_idx=q, bit_depth=bd, blocks=r["blocks"], block_txb=r["block_txb"], txb=r["txb"],
               coeffs=r["coeffs"], tile_cols=cols.value, tile_rows=rows.value, chroma_format=1 - ssx)
    return rec, r["recon"]


@pytest.mark.parametrize("W,H,q,seed,bd,edge", CASES)
def test_stream_parses_back_to_the_record(oracle, ref, W, H, q, seed, bd, edge):
    rec, _ = _record(oracle, W, H, q, seed, bd, edge)
    p = ref.parse(A.write_bitstream(rec, edge_filter=edge, avif=False))
    for k in ("blocks", "block_txb", "txb", "coeffs"):
        assert np.array_equal(p[k], rec[k]), k
    assert (p["base_q_idx"], p["width"], p["height"], p["bit_depth"], p["edge_filter"]) == (q, W, H, bd, edge)
    assert p["n_tiles"] == rec["tile_cols"] * rec["tile_rows"]


@pytest.mark.parametrize("W,H,q,seed,bd,edge", CASES)
def test_avif_decodes_to_the_reconstruction(oracle, W, H, q, seed, bd, edge):
    if not refcheck.dav1d_available():
        pytest.skip("libavif with dav1d (Pillow) is not available")
    rec, recon = _record(oracle, W, H, q, seed, bd, edge)
    d = refcheck.dav1d_decode_yuv(A.write_bitstream(rec, edge_filter=edge))
    for p in range(3):
        assert np.array_equal(d[p].astype(np.uint16), recon[p][:d[p].shape[0], :d[p].shape[1]]), "plane %d" % p


@pytest.mark.parametrize("W,H,q,seed,bd,edge", CASES444)
def test_444_stream_parses_back_and_decodes(oracle, ref, W, H, q, seed, bd, edge):
    rec, recon = _record(oracle, W, H, q, seed, bd, edge, ssx=0)
    obus = A.write_bitstream(rec, edge_filter=edge, avif=False)
    assert obus[2] >> 5 == 1                                      # seq_profile 1 (High) in the sequence header
    p = ref.parse(obus)
    for k in ("blocks", "block_txb", "txb", "coeffs"):
        assert np.array_equal(p[k], rec[k]), k
    if refcheck.dav1d_available():
        d = refcheck.dav1d_decode_yuv(A.write_bitstream(rec, edge_filter=edge))
        for pl in range(3):
            assert d[pl].shape == recon[pl][:H, :W].shape or d[pl].shape[1] == W   # full-resolution chroma
            assert np.array_equal(d[pl].astype(np.uint16), recon[pl][:d[pl].shape[0], :d[pl].shape[1]]), "plane %d" % pl


def _boxes(b, off=0, end=None):
    end = len(b) if end is None else end
    out = []
    while off < end:
        size, typ = struct.unpack(">I4s", b[off:off + 8])
        out.append((typ.decode(), off, size))
        off += size
    return out


def test_avif_container(oracle):
    rec, _ = _record(oracle, 64, 64, 120, 1, 8, 1)
    obus = A.write_bitstream(rec, avif=False)
    f = A.write_bitstream(rec)
    top = _boxes(f)
    assert [t for t, _, _ in top] == ["ftyp", "meta", "mdat"]
    assert f[8:12] == b"avif"
    _, moff, msize = top[2]
    assert f[moff + 8:moff + msize] == obus                      # mdat holds exactly the OBUs
    _, meta_off, meta_size = top[1]
    inner = [t for t, _, _ in _boxes(f, meta_off + 12, meta_off + meta_size)]
    assert inner == ["hdlr", "pitm", "iloc", "iinf", "iprp"]
    i = f.index(b"iloc")
    ext_off, ext_len = struct.unpack(">II", f[i + 4 + 4 + 2 + 2 + 2 + 2 + 2: i + 4 + 4 + 2 + 2 + 2 + 2 + 2 + 8])
    assert (ext_off, ext_len) == (moff + 8, len(obus))             # iloc points at the mdat payload
    assert obus[0] >> 3 == 1 and f[f.index(b"av1C") + 8:].startswith(obus[:obus[1] + 2])   # av1C carries the sequence header


def test_writer_refuses_an_inconsistent_record(oracle):
    rec, _ = _record(oracle, 64, 64, 120, 1, 8, 1)
    bad = dict(rec, tile_cols=2)
    with pytest.raises(ValueError):
        A.write_bitstream(bad)
    blocks = rec["blocks"].copy()
    blocks["mi_col"][0] += 2
    with pytest.raises(ValueError):
        A.write_bitstream(dict(rec, blocks=blocks))
