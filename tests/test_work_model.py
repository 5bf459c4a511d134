"""The work accounting behind bench.py's VALU roofline (forward-transform FLOPs of every RD evaluation), restated
from its definition: FLOPs of one evaluation by closed forms, and the exhaustive analysis's total as a closed form over
the shape grids (K candidates with DCT_DCT, then the set's other types for the winner).  The oracle's count on seeded
frames equals it; a miscounted pass or a wrong transform set would not."""
import av1gpu_amd as A

# AV1 intra transform sets (spec 5.11.48 get_tx_set / av1-spec tx sets): the types each tries, and whether a type's
# column / row pass is an identity
INTRA_1 = ("DCT_DCT", "ADST_DCT", "DCT_ADST", "ADST_ADST", "IDTX", "V_DCT", "H_DCT")
INTRA_2 = ("DCT_DCT", "ADST_DCT", "DCT_ADST", "ADST_ADST", "IDTX")
COL_IDT = {"IDTX", "H_DCT"}
ROW_IDT = {"IDTX", "V_DCT"}
SHAPES_WH = [(4, 4), (4, 8), (8, 4), (8, 8), (8, 16), (16, 8), (16, 16), (16, 32), (32, 16), (32, 32), (32, 64), (64, 32),
             (64, 64), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]


def flops_of(W, H, t):
    tw, th = min(W, 32), min(H, 32)
    col = W * th if t in COL_IDT else 2 * W * th * H
    row = th * tw if t in ROW_IDT else 2 * th * tw * W
    return col + row + th * tw


def intra_set(W, H):
    up = max(W, H)
    if up >= 32:
        return ("DCT_DCT",)
    return INTRA_2 if min(W, H) == 16 else INTRA_1


def exhaustive_analysis_flops(W, H, K=6):
    """Every block of every shape grid: K candidates with DCT_DCT, then the set's other types for the winner."""
    mic, mir = 2 * ((W + 7) >> 3), 2 * ((H + 7) >> 3)
    total = 0
    for w, h in SHAPES_WH:
        n = -(-mic * 4 // w) * -(-mir * 4 // h)
        per = K * flops_of(w, h, "DCT_DCT") + sum(flops_of(w, h, t) for t in intra_set(w, h) if t != "DCT_DCT")
        total += n * per
    return total


def test_flops_closed_forms():
    assert flops_of(4, 4, "DCT_DCT") == 2 * 64 + 2 * 64 + 16
    assert flops_of(4, 4, "IDTX") == 16 + 16 + 16
    assert flops_of(64, 64, "DCT_DCT") == 2 * 64 * 32 * 64 + 2 * 32 * 32 * 64 + 1024




def test_oracle_analysis_count_is_the_closed_form(oracle):
    for W, H, seed, q in ((264, 200, 6, 120), (72, 40, 5, 255), (8, 8, 4, 120)):
        r = oracle.encode(A.synth_image(seed, W, H), W, H, qindex=q, threads=4)
        assert r["flops"][0] == exhaustive_analysis_flops(W, H)
        assert r["flops"][1] > 0
