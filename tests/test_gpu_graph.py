"""Launch modes of the encode: analysis spread over 4 streams, the device part captured per batch shape as a hipGraph and
replayed (several replays in a row, with other eager encodes and a changed per-image qindex between them), the closed
loop's step schedule inside the graph, the learned pruning inside the graph, and the autotuner: every call gives the
eager single-stream records bit for bit."""
import numpy as np
import pytest

import av1gpu_amd as A
from av1gpu_amd import learn

pytestmark = pytest.mark.gpu

SPECS = [(200, 136, 80, 2001), (96, 64, 120, 2002), (264, 200, 160, 2003), (64, 130, 120, 2004)]


def _frames(qs=None):
    return [A.DeviceFrame(A.synth_image(s, W, H), W, H, qindex=(qs[i] if qs else q)) for i, (W, H, q, s) in enumerate(SPECS)]


def _eager(frames, **kw):
    e = A.GpEncoder().init(A.GpConfig(qindex=120, **kw))
    out = e.encode(frames)[0]
    e.close()
    return out


def _same(a, b):
    for x, y in zip(a, b):
        for k in ("blocks", "block_txb", "txb", "coeffs"):
            assert np.array_equal(getattr(x, k), getattr(y, k)), k
        for p in range(3):
            assert np.array_equal(x.recon[p], y.recon[p])


def test_analysis_streams_change_nothing():
    f = _frames()
    ref = _eager(f)
    e = A.GpEncoder().init(A.GpConfig(qindex=120, anStreams=4))
    out, t = e.encode(f)
    assert t.anStreams == 4
    _same(ref, out)
    e.close()


@pytest.mark.parametrize("schedule,streams", [(0, 1), (0, 4), (1, 4)])
def test_graph_replays_equal_eager(schedule, streams):
    f = _frames()
    ref = _eager(f, schedule=schedule)
    g = A.GpEncoder().init(A.GpConfig(qindex=120, graph=1, schedule=schedule, anStreams=streams))
    other = A.GpEncoder().init(A.GpConfig(qindex=120))
    extra = A.DeviceFrame(A.synth_image(2100, 136, 72), 136, 72)
    for rep in range(3):
        out, t = g.encode(f)
        assert t.graph == 1
        _same(ref, out)
        other.encode([extra])                       # eager work on the device between the replays
    # the same shape with other per-image qindex values: the replay reads this call's descriptors
    qs = [160, 80, 120, 200]
    f2 = _frames(qs)
    _same(_eager(f2, schedule=schedule), g.encode(f2)[0])
    # another shape: captured again
    f3 = f[:2]
    _same(_eager(f3, schedule=schedule), g.encode(f3)[0])
    for x in f + f2 + [extra]:
        x.close()
    g.close()
    other.close()


def test_graph_with_learned_pruning():
    f = _frames()
    p = learn.load_params()
    e = A.GpEncoder().init(A.GpConfig(qindex=120))
    learn.set_shortlist(e, p, 0.8)
    ref = e.encode(f)[0]
    e.close()
    g = A.GpEncoder().init(A.GpConfig(qindex=120, graph=1, anStreams=4))
    learn.set_shortlist(g, p, 0.8)
    for rep in range(2):
        _same(ref, g.encode(f)[0])
    learn.set_shortlist(g, None)                    # the captured graph held the network: captured again without it
    _same(_eager(f), g.encode(f)[0])
    g.close()
    for x in f:
        x.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_autotuner_tries_six_candidates_then_keeps_one(graph):
    f = _frames()
    ref = _eager(f)
    e = A.GpEncoder().init(A.GpConfig(qindex=120, tune=1, graph=graph))
    seen = set()
    for call in range(8):
        out, t = e.encode(f)
        _same(ref, out)
        key = (t.anStreams, t.schedule, t.queueOcc)
        if call < 6:
            assert t.tuned == 0
            seen.add(key)
        else:
            assert t.tuned == 1 and key in seen
    assert seen == {(s, 0, occ) for s in (1, 4) for occ in (1, 2)} | {(1, 1, 0), (4, 1, 0)}
    e.close()
    for x in f:
        x.close()
