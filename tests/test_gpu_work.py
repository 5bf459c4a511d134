"""Work accounting on the GPU (the forward-transform FLOPs behind bench.py's VALU roofline): the GPU's counts of the
analysis and of the closed loop equal the oracle's count of the evaluations it performs on the same input (so both run
the same set of evaluations), the exhaustive analysis equals the closed form over the shape grids, counting changes no
record, and the count is off (zero) by default."""
import numpy as np
import pytest

import av1gpu_amd as A
from test_work_model import exhaustive_analysis_flops

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H,seed,q,bd,yuv444", [(264, 200, 6, 120, 8, False), (136, 72, 50, 160, 10, True),
                                                   (130, 18, 8, 30, 8, False)])
def test_gpu_work_count_equals_oracle(oracle, W, H, seed, q, bd, yuv444):
    planes = A.synth_image(seed, W, H, ssx=0 if yuv444 else 1, bd=bd)
    f = A.DeviceFrame(planes, W, H, bd=bd, qindex=q, yuv444=yuv444)
    e = A.GpEncoder().init(A.GpConfig(qindex=q))
    g0, t0 = e.encode([f])
    assert t0.flopsAnalysis == 0 and t0.flopsRecon == 0
    e.set_work_count(True)
    g1, t1 = e.encode([f])
    e.close()
    f.close()
    for k in ("blocks", "txb", "coeffs"):
        assert np.array_equal(getattr(g0[0], k), getattr(g1[0], k)), k
    r = oracle.encode(planes, W, H, ssx=0 if yuv444 else 1, bd=bd, qindex=q)
    assert (t1.flopsAnalysis, t1.flopsRecon) == r["flops"]
    assert t1.flopsAnalysis == exhaustive_analysis_flops(W, H)
    assert t1.flopsRecon > 0
