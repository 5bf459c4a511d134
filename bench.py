"""Benchmark of the intra RD core on one MI355X.

Headline (value): the batched+learned path.  A step is one encode call (analysis with the learned pruning network ->
partition DP -> closed-loop reconstruction -> pack -> records on the host) over the mixed-size batch: 256 synthetic
8-bit 4:2:0 images of 0.18-24 MP, aspect 0.5-2, qindex 80/120/160 (av1gpu_amd.batch_spec, seeds 1000..1255), planes
resident in device memory, the shipped network (av1gpu_amd/shortlist_params.npy) at confidence 0.8.  Beside it:
  plain_single_image   the plain exhaustive search on one 4000 x 3000 image (seed 1000) at qindex 120, the path the
                       ratio vs_plain_single_image divides by;
  plain_image_by_image the plain exhaustive search over the same 256-image batch, one encode call per image (the
                       ratio vs_plain_image_by_image divides by it);
  batched_exhaustive   the same batch without the network;
  bd_rate              BD-rate (PSNR-Y, bits of this build's own writer) of the learned search against the exhaustive
                       one on 12 batch images stratified by size (0.18-24 MP), qindex 60/100/140/180;
  learned_10bit_444    the same batch at 10-bit 4:4:4 (learned and exhaustive, with its BD-rate);
  serving              Poisson arrivals into the continuous-batching server on the batched+learned path, p50 / p99, at
                       50 req/s (serving_rates: 50, 200 and 400 req/s);
  cpu_baseline         the CPU restatement (oracle/liborc.so): warm-up + median of 5 on the 512 x 512 configuration
                       (1 thread and all threads) and on the whole 4000 x 3000 image (all threads, up to 16).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--batch 256] [--serve 64] [--conf 0.8]

roofline: the dominant kernel of the headline step, the closed loop k_recon_queue (one persistent launch per encode):
its algorithmic HBM bytes (source planes read once + reconstruction written once + 2 bytes per coefficient + 16 bytes
per block and tx-block record) over its launch duration, timed with HIP events on the encode stream.
roofline_valu: the same launch (and the analysis stage) against the FP32 vector peak, from the algorithmic
forward-transform FLOPs of every RD evaluation, counted by the encoder's work accounting in an untimed encode.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MP/s at qindex 120, batched+learned vs plain, at <=+1% BD-rate; p99 latency"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 157.3   # MI355X dense FP32 vector peak
FLOPS_DEF = ("forward-transform FLOPs of every RD evaluation the search performs (2 per multiply-add of the column and "
             "row passes, 1 per multiply of identity kinds and of the row scale; 64-point sizes keep 32 outputs), counted "
             "by the encoder's work accounting in an untimed encode of the same workload")


def valu_roofline(kernel, flops, ms):
    """FP32 VALU roofline of a kernel (or stage) from its algorithmic forward-transform FLOPs per launch and its launch
    duration."""
    achieved = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    return {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 4), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / VALU_PEAK_TFLOPS, "flops_per_launch": int(flops), "per_launch_ms": round(ms, 4)}


def counted_flops(enc, frames):
    """(analysis, closed loop) forward-transform FLOPs of one encode of frames (an untimed, counting call)."""
    enc.set_work_count(True)
    _, t = enc.encode(frames, want_recon=False)
    enc.set_work_count(False)
    return t.flopsAnalysis, t.flopsRecon


def measured_traffic(kernel, name="counters_batched_learned_256.md"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE of the same
    workload, calibrated against a copy kernel of known size), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic_bytes_per_launch" and w[1] == kernel:
                    return int(w[2])
    except OSError:
        pass
    return None


def alg_bytes(recs, bps=1):
    """Algorithmic HBM bytes of the closed loop over records: planes read + reconstruction written (4:2:0), 2 bytes
    per coefficient, 16 per block and tx-block entry."""
    return sum(r.width * r.height * bps * 1.5 * 2 + 2 * len(r.coeffs) + 16 * (len(r.blocks) + len(r.txb)) for r in recs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256, help="images of the mixed-size batch")
    ap.add_argument("--conf", type=float, default=0.8, help="confidence of the learned pruning")
    ap.add_argument("--width", type=int, default=4000)
    ap.add_argument("--height", type=int, default=3000)
    ap.add_argument("--plain-steps", type=int, default=3, help="timed encodes of the plain leg (0 = skip it)")
    ap.add_argument("--no-exhaustive", action="store_true", help="skip the batched exhaustive leg")
    ap.add_argument("--no-by-image", action="store_true", help="skip the plain image-by-image leg over the batch")
    ap.add_argument("--plain-only", action="store_true", help="run the plain leg alone (profiling)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-bd", action="store_true", help="skip the BD-rate leg")
    ap.add_argument("--schedule", type=int, default=0, help="closed loop: 0 persistent queue, 1 one launch per wavefront step")
    ap.add_argument("--no-444", action="store_true", help="skip the 10-bit 4:4:4 batched+learned leg")
    ap.add_argument("--no-modes", action="store_true", help="skip the launch-mode leg (streams / hipGraph / autotuner)")
    ap.add_argument("--serve", type=int, default=64, help="requests of the serving leg (0 = skip it)")
    ap.add_argument("--rate", default="50,200,400", help="serving leg: Poisson arrival rates, requests per second (comma list)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path shards by image / tile with no collective); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import av1gpu_amd as A
    from av1gpu_amd import learn

    plain, planes = plain_leg(A, a.width, a.height, 120, a.schedule, a.plain_steps) if a.plain_steps > 0 else (None, None)
    if a.plain_only:
        print(json.dumps({"plain_single_image": plain}), flush=True)
        return
    params = learn.load_params()

    from concurrent.futures import ThreadPoolExecutor
    spec = A.batch_spec(a.batch)
    with ThreadPoolExecutor(16) as ex:
        bplanes = list(ex.map(lambda s: A.synth_image(s[3], s[0], s[1]), spec))
    frames = [A.DeviceFrame(p, W, H, qindex=q) for p, (W, H, q, _) in zip(bplanes, spec)]
    mp = sum(W * H for W, H, _, _ in spec) / 1e6

    enc = A.GpEncoder().init(A.GpConfig(qindex=120, schedule=a.schedule))
    learn.set_shortlist(enc, params, a.conf)
    for _ in range(a.warmup):
        enc.encode(frames, want_recon=False)
    A.lib().av1mi_device_sync()
    stage = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        out, t = enc.encode(frames, want_recon=False)
        stage.append(t)
        ab = alg_bytes(out)
        out = None          # the records are consumed: their host buffers go back before the next call
    A.lib().av1mi_device_sync()
    elapsed = time.perf_counter() - t0
    fl_an, fl_rc = counted_flops(enc, frames)
    enc.close()
    value = mp * a.steps / elapsed
    avg = lambda k: round(sum(getattr(s, k) for s in stage) / len(stage), 3)
    last = stage[-1]
    launches = 1 if last.schedule == 0 else max(last.steps, 1)
    kernel = "k_recon_queue" if last.schedule == 0 else "k_recon"
    per_launch_ms = avg("recon") / launches
    achieved = (ab / launches) / (per_launch_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "batched+learned: %d synthetic 8-bit 4:2:0 images, 0.18-24 MP log-uniform, aspect 0.5-2, "
                               "qindex 80/120/160 (seeds 1000..%d), one encode call per step, learned pruning at "
                               "confidence %g" % (a.batch, 1000 + a.batch - 1, a.conf),
                   "images": a.batch, "megapixels": round(mp, 3), "confidence": a.conf, "parallelism": "single GPU"},
        "vs_plain_single_image": round(value / plain["value"], 3) if plain else None,
        "stages_ms": {k: avg(k) for k in ("analysis", "dp", "recon", "pack", "download", "total")},
        "record": {"sbs": last.sbs, "blocks": last.blocks, "device_bytes": last.deviceBytes},
        "schedule": {"closed_loop": "persistent queue" if last.schedule == 0 else "launch per wavefront step",
                     "workgroups": last.queueBlocks},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kernel) if launches == 1 else None,
                     "traffic_source": "profiles/counters_batched_learned_256.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, "
                                       "factors from a 1 GiB copy)",
                     "per_launch_ms": round(per_launch_ms, 4), "alg_bytes_per_launch": round(ab / launches)},
        "roofline_valu": {"definition": FLOPS_DEF,
                          "closed_loop": valu_roofline(kernel, fl_rc / launches, per_launch_ms),
                          "analysis_stage": valu_roofline("k_analysis (19 shape launches + the pruning network)", fl_an,
                                                          avg("analysis"))},
        "plain_single_image": plain,
    }
    print("headline done: %.1f MP/s" % value, file=sys.stderr, flush=True)
    if not a.no_by_image:
        line["plain_image_by_image"] = plain_by_image_leg(A, frames, mp, a.schedule)
        line["vs_plain_image_by_image"] = round(value / line["plain_image_by_image"]["value"], 3)
    if not a.no_exhaustive:
        line["batched_exhaustive"] = exhaustive_leg(A, frames, mp, a.schedule, value)
    for f in frames:
        f.close()
    if not a.no_bd:
        line["bd_rate"] = bd_leg(A, learn, params, a.conf, spec, bplanes)
    del bplanes
    if not a.no_modes:
        line["launch_modes"] = modes_leg(A, learn, params, a.conf)
    if not a.no_444:
        line["learned_10bit_444"] = leg_444(A, learn, params, a.conf, a.batch, a.schedule)
    if a.serve > 0:
        rates = [float(r) for r in str(a.rate).split(",") if r]
        legs = [serving_leg(A, a.serve, r, (params, a.conf)) for r in rates]
        line["serving"] = legs[0]
        line["serving_rates"] = [{k: v[k] for k in ("requests", "value", "p50_ms", "p99_ms", "batches")} | {"rate": r}
                                 for r, v in zip(rates, legs)]
    if not a.no_cpu and planes is not None:
        line["cpu_baseline"] = cpu_baseline(A, planes, a.width, a.height, 120)
    print(json.dumps(line), flush=True)


def plain_leg(A, W, H, qindex, schedule, steps):
    """The plain exhaustive search on one W x H image, steps timed encodes after one warm-up (records to the host, the
    reconstruction kept on the device, as in the headline)."""
    planes = A.synth_image(1000, W, H)
    frame = A.DeviceFrame(planes, W, H, qindex=qindex)
    enc = A.GpEncoder().init(A.GpConfig(qindex=qindex, schedule=schedule))
    enc.encode([frame], want_recon=False)
    A.lib().av1mi_device_sync()
    stage = []
    t0 = time.perf_counter()
    for _ in range(steps):
        out, t = enc.encode([frame], want_recon=False)
        stage.append(t)
    A.lib().av1mi_device_sync()
    elapsed = time.perf_counter() - t0
    fl_an, fl_rc = counted_flops(enc, [frame])
    enc.close()
    frame.close()
    avg = lambda k: round(sum(getattr(s, k) for s in stage) / len(stage), 3)
    rec, last = out[0], stage[-1]
    launches = 1 if last.schedule == 0 else max(last.steps, 1)
    ab = alg_bytes([rec])
    achieved = (ab / launches) / (avg("recon") / launches * 1e-3) / 1e9
    res = {"workload": "plain exhaustive intra search, one %dx%d 8-bit 4:2:0 synthetic image (seed 1000), qindex %d"
                       % (W, H, qindex), "value": round(W * H / 1e6 * steps / elapsed, 4), "unit": "MP/s", "steps": steps,
           "ms_per_step": round(1e3 * elapsed / steps, 3),
           "stages_ms": {k: avg(k) for k in ("analysis", "dp", "recon", "pack", "download", "total")},
           "record": {"blocks": len(rec.blocks), "txb": len(rec.txb), "coeffs": len(rec.coeffs), "wavefront_steps": last.steps,
                      "sbs": last.sbs},
           "roofline": {"bound": "hbm", "kernel": "k_recon_queue" if last.schedule == 0 else "k_recon",
                        "achieved": round(achieved, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "per_launch_ms": round(avg("recon") / launches, 4)},
           "roofline_valu": {"closed_loop": valu_roofline("k_recon_queue" if last.schedule == 0 else "k_recon", fl_rc / launches,
                                                          avg("recon") / launches),
                             "analysis_stage": valu_roofline("k_analysis (19 shape launches)", fl_an, avg("analysis"))}}
    print("plain leg done: %.1f MP/s" % res["value"], file=sys.stderr, flush=True)
    return res, planes


def plain_by_image_leg(A, frames, mp, schedule):
    """The plain single-image path over the same batch: the exhaustive search, one encode call per image (each at its
    own qindex), in batch order, records to the host, after one untimed warm-up call."""
    enc = A.GpEncoder().init(A.GpConfig(qindex=120, schedule=schedule))
    enc.encode(frames[:1], want_recon=False)
    A.lib().av1mi_device_sync()
    t0 = time.perf_counter()
    for f in frames:
        enc.encode([f], want_recon=False)
    A.lib().av1mi_device_sync()
    dt = time.perf_counter() - t0
    enc.close()
    res = {"workload": "plain exhaustive search, one encode call per image over the headline batch (%d images, %.1f MP)"
                       % (len(frames), mp), "value": round(mp / dt, 4), "unit": "MP/s", "seconds": round(dt, 3)}
    print("image-by-image leg done: %.1f MP/s" % res["value"], file=sys.stderr, flush=True)
    return res


def exhaustive_leg(A, frames, mp, schedule, learned_value):
    """The same batch with the exhaustive search: one untimed pass, one timed."""
    enc = A.GpEncoder().init(A.GpConfig(qindex=120, schedule=schedule))
    enc.encode(frames[:8], want_recon=False)
    A.lib().av1mi_device_sync()
    t0 = time.perf_counter()
    _, t = enc.encode(frames, want_recon=False)
    A.lib().av1mi_device_sync()
    dt = time.perf_counter() - t0
    enc.close()
    res = {"value": round(mp / dt, 4), "unit": "MP/s", "seconds": round(dt, 3),
           "learned_speedup": round(learned_value / (mp / dt), 3),
           "stages_ms": {"analysis": round(t.analysis, 3), "dp": round(t.dp, 3), "recon": round(t.recon, 3),
                         "pack": round(t.pack, 3), "download": round(t.download, 3)}, "blocks": t.blocks}
    print("exhaustive batch done: %.1f MP/s" % res["value"], file=sys.stderr, flush=True)
    return res


def stratified(spec, n):
    """n batch images spread evenly over the size order (the smallest and the largest included)."""
    order = sorted(range(len(spec)), key=lambda i: spec[i][0] * spec[i][1])
    return sorted(set(order[int(round(k * (len(order) - 1) / (n - 1)))] for k in range(n)))


def bd_leg(A, learn, params, conf, spec, bplanes, n=12, qs=(60, 100, 140, 180)):
    """BD-rate of the learned search against the exhaustive one on n batch images stratified by size (0.18-24 MP)."""
    idx = stratified(spec, n)
    images = [(bplanes[i], spec[i][0], spec[i][1]) for i in idx]
    curves = {}
    for name, c in (("exhaustive", None), ("learned", conf)):
        enc = A.GpEncoder().init(A.GpConfig(qindex=120))
        if c is not None:
            learn.set_shortlist(enc, params, c)
        bits, psnr, secs = learn.rd_curve(enc, images, qs)
        enc.close()
        curves[name] = {"bits": bits, "psnr_y": [round(v, 4) for v in psnr], "seconds": round(secs, 3)}
    bd = learn.bd_rate(curves["exhaustive"]["bits"], curves["exhaustive"]["psnr_y"], curves["learned"]["bits"],
                       curves["learned"]["psnr_y"])
    print("bd leg done: %.3f %%" % bd, file=sys.stderr, flush=True)
    return {"bd_rate_psnr_y_pct": round(float(bd), 3), "images": [spec[i][3] for i in idx],
            "sizes_mp": [round(W * H / 1e6, 2) for _, W, H in images],
            "megapixels": round(sum(W * H for _, W, H in images) / 1e6, 3), "qindex": list(qs), "curves": curves}


def modes_leg(A, learn, params, conf, n=16, W=512, H=512, calls=10):
    """Launch modes on a workload where launches matter: batches of n small images (the same batch shape every call,
    learned pruning on): eager on one stream, analysis over 4 streams, hipGraph replay, and the autotuner (its 6 trial
    calls in the warm-up).  Each mode: untimed calls until warm, then `calls` timed calls."""
    frames = [A.DeviceFrame(A.synth_image(3000 + i, W, H), W, H, qindex=120) for i in range(n)]
    mp = n * W * H / 1e6
    res = {"workload": "%d images of %dx%d 8-bit 4:2:0 (seeds 3000..%d), qindex 120, learned pruning, the same batch every "
                       "call" % (n, W, H, 3000 + n - 1), "megapixels_per_call": mp}
    for name, kw, warm in (("eager_1_stream", {}, 2), ("eager_4_streams", {"anStreams": 4}, 2),
                           ("graph_4_streams", {"anStreams": 4, "graph": 1}, 2), ("tuned_graph", {"tune": 1, "graph": 1}, 7)):
        enc = A.GpEncoder().init(A.GpConfig(qindex=120, **kw))
        learn.set_shortlist(enc, params, conf)
        for _ in range(warm):
            enc.encode(frames, want_recon=False)
        A.lib().av1mi_device_sync()
        t0 = time.perf_counter()
        for _ in range(calls):
            out, t = enc.encode(frames, want_recon=False)
        A.lib().av1mi_device_sync()
        dt = (time.perf_counter() - t0) / calls
        res[name] = {"ms_per_call": round(dt * 1e3, 3), "value": round(mp / dt, 3), "unit": "MP/s",
                     "device_ms": round(t.analysis + t.dp + t.recon, 3), "anStreams": t.anStreams, "schedule": t.schedule,
                     "queueOcc": t.queueOcc, "graph": t.graph, "tuned": t.tuned}
        enc.close()
    for f in frames:
        f.close()
    print("modes leg done: %s" % {k: v["value"] for k, v in res.items() if isinstance(v, dict)}, file=sys.stderr, flush=True)
    return res


def leg_444(A, learn, params, conf, n, schedule, qs=(60, 100, 140, 180)):
    """The same batch at 10-bit 4:4:4 (generator scaled to [0, 1023], chroma at full resolution): one timed
    batched+learned encode after a warm-up on 8 images, one exhaustive encode, and the BD-rate (PSNR-Y) of the learned
    search against the exhaustive one on the first 8 images of at most 2 MP."""
    from concurrent.futures import ThreadPoolExecutor
    spec = A.batch_spec(n)
    with ThreadPoolExecutor(16) as ex:
        planes = list(ex.map(lambda s: A.synth_image(s[3], s[0], s[1], ssx=0, bd=10), spec))
    frames = [A.DeviceFrame(p, W, H, bd=10, qindex=q, yuv444=True) for p, (W, H, q, _) in zip(planes, spec)]
    mp = sum(W * H for W, H, _, _ in spec) / 1e6
    res = {"workload": "%d synthetic 10-bit 4:4:4 images, same sizes / qindex / seeds as the headline batch" % n,
           "megapixels": round(mp, 3), "unit": "MP/s"}
    for name, c in (("learned", conf), ("exhaustive", None)):
        enc = A.GpEncoder().init(A.GpConfig(qindex=120, schedule=schedule))
        if c is not None:
            learn.set_shortlist(enc, params, c)
        enc.encode(frames[:8], want_recon=False)
        A.lib().av1mi_device_sync()
        t0 = time.perf_counter()
        out, t = enc.encode(frames, want_recon=False)
        A.lib().av1mi_device_sync()
        dt = time.perf_counter() - t0
        out = None
        enc.close()
        res[name] = {"value": round(mp / dt, 4), "seconds": round(dt, 3),
                     "stages_ms": {"analysis": round(t.analysis, 3), "recon": round(t.recon, 3), "download": round(t.download, 3)}}
    for f in frames:
        f.close()
    res["learned_speedup"] = round(res["learned"]["value"] / res["exhaustive"]["value"], 3)
    idx = [i for i, (W, H, _, _) in enumerate(spec) if W * H <= 2e6][:8]
    images = [(planes[i], spec[i][0], spec[i][1]) for i in idx]
    curves = {}
    for name, c in (("exhaustive", None), ("learned", conf)):
        enc = A.GpEncoder().init(A.GpConfig(qindex=120))
        if c is not None:
            learn.set_shortlist(enc, params, c)
        bits, psnr, _ = learn.rd_curve(enc, images, qs, bd=10, yuv444=True)
        enc.close()
        curves[name] = {"bits": bits, "psnr_y": [round(v, 4) for v in psnr]}
    res["bd_rate_psnr_y_pct"] = round(float(learn.bd_rate(curves["exhaustive"]["bits"], curves["exhaustive"]["psnr_y"],
                                                          curves["learned"]["bits"], curves["learned"]["psnr_y"])), 3)
    res["bd_images"] = [spec[i][3] for i in idx]
    print("4:4:4 leg done: %.1f MP/s learned, %.1f exhaustive, BD %.3f %%" % (res["learned"]["value"], res["exhaustive"]["value"],
                                                                             res["bd_rate_psnr_y_pct"]), file=sys.stderr, flush=True)
    return res


def serving_leg(A, n, rate, shortlist):
    """n single-image requests (serve.serving_spec: 0.2-24 MP, 1 % at 24 MP, qindex 120) arriving as a Poisson
    process at `rate` per second into the continuous-batching server on the batched+learned path; latency =
    submission to result on the host."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    from av1gpu_amd import serve
    spec = serve.serving_spec(n)
    with ThreadPoolExecutor(16) as ex:
        planes = list(ex.map(lambda s: A.synth_image(s[3], s[0], s[1]), spec))
    images = [(p, W, H, q) for p, (W, H, q, _) in zip(planes, spec)]
    srv = serve.Server(workers=4, hbm_budget=96 << 30, max_batch=16, shortlist=shortlist)
    serve.poisson_run(srv, images[:4], rate)          # warm-up
    nb0 = len(srv.batches)
    reqs, wall = serve.poisson_run(srv, images, rate)
    srv.close()
    lat = np.array([r.latency for r in reqs]) * 1e3
    mp = sum(W * H for W, H, _, _ in spec) / 1e6
    return {"workload": "%d requests, 8-bit 4:2:0, 0.2-24 MP log-uniform with 1%% at 24 MP, qindex 120, Poisson %.0f/s"
                        % (n, rate), "requests": n, "megapixels": round(mp, 3), "value": round(mp / wall, 4), "unit": "MP/s",
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "workers": 4, "batches": len(srv.batches) - nb0, "learned_confidence": shortlist[1]}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A, planes, W, H, qindex, runs=5):
    """The oracle (the CPU restatement of the same exhaustive search, oracle/liborc.so) on this host: one warm-up, then
    the median of `runs` timed encodes, for the 512 x 512 CPU configuration on one thread and on all threads (up to 16)
    and for the whole 4000 x 3000 benchmark image on all threads (its value is `value`)."""
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refcheck
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1))
    o = refcheck.Oracle()

    def timed(pl, w, h, nt):
        o.encode(pl, w, h, qindex=qindex, threads=nt)
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            o.encode(pl, w, h, qindex=qindex, threads=nt)
            ts.append(time.perf_counter() - t0)
        med = statistics.median(ts)
        return {"threads": nt, "median_s": round(med, 4), "runs_s": [round(t, 4) for t in ts], "mp_per_s": round(w * h / 1e6 / med, 5)}

    small = A.synth_image(1000, 512, 512)
    legs = {"512x512_1_thread": timed(small, 512, 512, 1), "512x512_all_threads": timed(small, 512, 512, threads),
            "%dx%d_all_threads" % (W, H): timed(planes, W, H, threads)}
    big = legs["%dx%d_all_threads" % (W, H)]
    print("cpu leg done: %.3f MP/s on %d threads" % (big["mp_per_s"], threads), file=sys.stderr, flush=True)
    return {"value": big["mp_per_s"], "unit": "MP/s", "cores": threads, "kind": "port",
            "sample": "oracle/liborc.so, the same exhaustive search on the whole %dx%d benchmark image (seed 1000, qindex %d), "
                      "%d threads, median of %d after a warm-up" % (W, H, qindex, threads, runs),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "legs": legs}


if __name__ == "__main__":
    main()
