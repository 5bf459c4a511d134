"""Benchmark of the intra RD core on one MI355X.

A step is one encode (analysis -> partition DP -> closed-loop reconstruction -> pack -> records and reconstruction on
the host) of the workload with its input planes already resident in device memory.  Default workload: the plain
exhaustive search on one 4000 x 3000 8-bit 4:2:0 synthetic image (seed 1000) at qindex 120.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

roofline: the dominant kernel's (the closed loop's) algorithmic HBM bytes (source planes read once + reconstruction
written once + 2 bytes per coefficient + 16 bytes per block and tx-block record) per launch over its average launch
duration, timed with HIP events on the encode stream.  cpu_baseline: the CPU restatement (oracle/liborc.so) on a bounded crop of the
same image with all host threads.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=4000)
    ap.add_argument("--height", type=int, default=3000)
    ap.add_argument("--qindex", type=int, default=120)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--schedule", type=int, default=0, help="closed loop: 0 persistent queue, 1 one launch per wavefront step")
    ap.add_argument("--batch", type=int, default=256, help="images of the mixed-size batch leg (0 = skip it)")
    ap.add_argument("--serve", type=int, default=64, help="requests of the serving leg (0 = skip it)")
    ap.add_argument("--rate", type=float, default=50.0, help="serving leg: Poisson arrival rate, requests per second")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path shards by image / tile with no collective); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import av1gpu_amd as A

    W, H = a.width, a.height
    planes = A.synth_image(1000, W, H)
    frame = A.DeviceFrame(planes, W, H, qindex=a.qindex)
    enc = A.GpEncoder().init(A.GpConfig(qindex=a.qindex, schedule=a.schedule))
    for _ in range(a.warmup):
        enc.encode([frame], want_recon=True)
    A.lib().av1mi_device_sync()
    times, stage = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s0 = time.perf_counter()
        out, t = enc.encode([frame], want_recon=True)
        times.append(time.perf_counter() - s0)
        stage.append(t)
    A.lib().av1mi_device_sync()
    elapsed = time.perf_counter() - t0
    mp = W * H / 1e6
    value = mp * a.steps / elapsed
    rec = out[0]
    n_coeffs, n_blocks, n_txb = len(rec.coeffs), len(rec.blocks), len(rec.txb)
    alg_bytes = W * H * 1 * 1.5 * 2 + 2 * n_coeffs + 16 * (n_blocks + n_txb)
    recon_ms = sum(s.recon for s in stage) / len(stage)
    analysis_ms = sum(s.analysis for s in stage) / len(stage)
    steps = stage[-1].steps
    # dominant kernel: the closed loop, one persistent launch per encode (queue schedule) or one per wavefront step
    launches = 1 if stage[-1].schedule == 0 else max(steps, 1)
    kernel = "k_recon_queue" if stage[-1].schedule == 0 else "k_recon"
    per_launch_ms = recon_ms / launches
    achieved = (alg_bytes / launches) / (per_launch_ms * 1e-3) / 1e9
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
        "config": {"workload": "plain exhaustive intra search, one %dx%d 8-bit 4:2:0 synthetic image (seed 1000), "
                               "qindex %d" % (W, H, a.qindex), "width": W, "height": H, "qindex": a.qindex,
                   "parallelism": "single GPU"},
        "stages_ms": {"analysis": round(analysis_ms, 3), "dp": round(sum(s.dp for s in stage) / len(stage), 3),
                      "recon": round(recon_ms, 3), "pack": round(sum(s.pack for s in stage) / len(stage), 3),
                      "download": round(sum(s.download for s in stage) / len(stage), 3),
                      "total_events": round(sum(s.total for s in stage) / len(stage), 3)},
        "record": {"blocks": n_blocks, "txb": n_txb, "coeffs": n_coeffs, "wavefront_steps": steps, "sbs": stage[-1].sbs},
        "schedule": {"closed_loop": "persistent queue" if stage[-1].schedule == 0 else "launch per wavefront step",
                     "workgroups": stage[-1].queueBlocks},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "per_launch_ms": round(per_launch_ms, 4), "alg_bytes_per_launch": round(alg_bytes / launches)},
    }
    frame.close()
    if a.batch > 0:
        line["batched"] = batch_leg(A, a.batch, a.schedule, value)
    if a.serve > 0:
        line["serving"] = serving_leg(A, a.serve, a.rate)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(planes, W, H, a.qindex)
    print(json.dumps(line), flush=True)
    enc.close()


def batch_leg(A, n, schedule, plain_value):
    """The mixed-size batch (A.batch_spec: n images, 0.18-24 MP, qindex 80/120/160) in one encode call, inputs
    resident in device memory; one warm-up call on the first 8 images, then one timed call on all n."""
    from concurrent.futures import ThreadPoolExecutor
    spec = A.batch_spec(n)
    with ThreadPoolExecutor(16) as ex:
        planes = list(ex.map(lambda s: A.synth_image(s[3], s[0], s[1]), spec))
    frames = [A.DeviceFrame(p, W, H, qindex=q) for p, (W, H, q, _) in zip(planes, spec)]
    del planes
    enc = A.GpEncoder().init(A.GpConfig(qindex=120, schedule=schedule))
    enc.encode(frames[:8], want_recon=True)
    A.lib().av1mi_device_sync()
    t0 = time.perf_counter()
    out, t = enc.encode(frames, want_recon=True)
    A.lib().av1mi_device_sync()
    dt = time.perf_counter() - t0
    mp = sum(W * H for W, H, _, _ in spec) / 1e6
    res = {"workload": "%d synthetic 8-bit 4:2:0 images, 0.18-24 MP log-uniform, aspect 0.5-2, qindex 80/120/160, "
                       "one encode call (seeds 1000..%d)" % (n, 1000 + n - 1),
           "images": n, "megapixels": round(mp, 3), "value": round(mp / dt, 4), "unit": "MP/s", "seconds": round(dt, 3),
           "vs_plain_single_image": round(mp / dt / plain_value, 3),
           "stages_ms": {"analysis": round(t.analysis, 3), "dp": round(t.dp, 3), "recon": round(t.recon, 3),
                         "pack": round(t.pack, 3), "download": round(t.download, 3)},
           "sbs": t.sbs, "blocks": t.blocks, "device_bytes": t.deviceBytes}
    for f in frames:
        f.close()
    enc.close()
    return res


def serving_leg(A, n, rate):
    """n single-image requests (serve.serving_spec: 0.2-24 MP, 1 % at 24 MP, qindex 120) arriving as a Poisson
    process at `rate` per second into the continuous-batching server; latency = submission to result on the host."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    from av1gpu_amd import serve
    spec = serve.serving_spec(n)
    with ThreadPoolExecutor(16) as ex:
        planes = list(ex.map(lambda s: A.synth_image(s[3], s[0], s[1]), spec))
    images = [(p, W, H, q) for p, (W, H, q, _) in zip(planes, spec)]
    srv = serve.Server(workers=4, hbm_budget=96 << 30, max_batch=16)
    serve.poisson_run(srv, images[:4], rate)          # warm-up
    nb0 = len(srv.batches)
    reqs, wall = serve.poisson_run(srv, images, rate)
    srv.close()
    lat = np.array([r.latency for r in reqs]) * 1e3
    mp = sum(W * H for W, H, _, _ in spec) / 1e6
    return {"workload": "%d requests, 8-bit 4:2:0, 0.2-24 MP log-uniform with 1%% at 24 MP, qindex 120, Poisson %.0f/s"
                        % (n, rate), "requests": n, "megapixels": round(mp, 3), "value": round(mp / wall, 4), "unit": "MP/s",
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "workers": 4, "batches": len(srv.batches) - nb0}


def cpu_baseline(planes, W, H, qindex):
    """The oracle (the CPU restatement of the same search) on a 1024 x 768 crop of the same image, all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refcheck
    cw, ch = min(1024, W), min(768, H)
    crop = [planes[0][:ch, :cw], planes[1][:(ch + 1) // 2, :(cw + 1) // 2], planes[2][:(ch + 1) // 2, :(cw + 1) // 2]]
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1))
    o = refcheck.Oracle()
    t0 = time.perf_counter()
    o.encode(crop, cw, ch, qindex=qindex, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(cw * ch / 1e6 / dt, 5), "unit": "MP/s", "cores": threads, "kind": "port",
            "sample": "oracle/liborc.so, the same search on the top-left %dx%d crop of the benchmark image, %d threads, "
                      "%.2f s" % (cw, ch, threads, dt)}


if __name__ == "__main__":
    main()
