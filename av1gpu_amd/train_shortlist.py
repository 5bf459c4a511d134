"""Train the learned pruning network on the GPU and evaluate it against the exhaustive search.

  python -m av1gpu_amd.train_shortlist [--sbs 2000000] [--steps 60000] [--train-budget SECONDS] [--out av1gpu_amd/shortlist_params.npy]

1. Labels: seeded synthetic images (sizes log-uniform over 0.2-4 MP, aspect 0.5-2, qindex 80/120/160, seeds from
   20000) are encoded with the exhaustive search; every superblock's leaves give its 3 x 85 node labels and 13 mode
   labels (learn.sb_labels; with --rd-tau the node labels come from the partition DP's costs, learn.rd_labels), its
   64x64 luma tile is the input.
2. Training: Adam (lr cosine-decayed from --lr), batches of --batch superblocks, on the library's own kernels.
3. Evaluation on held-out images (seeds from 30000) at qindex 60/100/140/180: BD-rate (PSNR-Y, bits of this build's
   own bitstream writer) of the pruned search against the exhaustive one, and the share of analysis work pruned.
Writes the parameters (float32 .npy) and a JSON summary next to them.
"""
import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import av1gpu_amd as A
from av1gpu_amd import learn as Ln


def image_spec(n, seed, mp_lo=0.2, mp_hi=4.0, qs=(80, 120, 160)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        mp = math.exp(rng.uniform(math.log(mp_lo), math.log(mp_hi)))
        a = rng.uniform(0.5, 2.0)
        W = max(64, int(round(math.sqrt(mp * 1e6 * a) / 8.0)) * 8)
        H = max(64, int(round(math.sqrt(mp * 1e6 / a) / 8.0)) * 8)
        out.append((W, H, int(rng.choice(qs)), seed + i))
    return out


def labelled_data(n_sbs, seed=20000, log=print, rd_tau=None, rd_tau_root=None):
    """Tiles, qindex, labels and masks of about n_sbs superblocks of exhaustively encoded images.  rd_tau: node labels
    from the DP's relative cost increases (learn.rd_labels) instead of the leaves' use."""
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
    if rd_tau is not None:
        Ln.dp_export(enc, True)
    tiles, qs, labs, msks = [], [], [], []
    total, k = 0, 0
    t0 = time.time()
    while total < n_sbs:
        spec = image_spec(32, seed + k)
        k += 32
        with ThreadPoolExecutor(16) as ex:
            planes = list(ex.map(lambda s: A.synth_image(s[3], s[0], s[1]), spec))
        frames = [A.DeviceFrame(p, W, H, qindex=q) for p, (W, H, q, _) in zip(planes, spec)]
        out, _ = enc.encode(frames, want_recon=False)
        rel = Ln.dp_export(enc, None) if rd_tau is not None else None
        for f in frames:
            f.close()
        s0 = 0
        for p, (W, H, q, _), o in zip(planes, spec, out):
            lab, msk = Ln.sb_labels(o.blocks, W, H)
            if rel is not None:
                lab = Ln.rd_labels(rel[s0:s0 + len(lab)], lab, rd_tau, rd_tau_root)
                s0 += len(lab)
            t = Ln.sb_tiles(p[0], W, H)
            tiles.append(t); qs.append(np.full(len(t), q, np.uint8)); labs.append(lab); msks.append(msk)
            total += len(t)
        log("data: %d superblocks from %d images (%.0f s)" % (total, k, time.time() - t0))
    enc.close()
    return np.concatenate(tiles), np.concatenate(qs), np.concatenate(labs), np.concatenate(msks)


def train(params, data, steps, batch, lr, log=print, seed=7, budget=None):
    """budget (seconds): after 20 steps the step count is cut to what fits the budget at the measured step time."""
    tiles, q, lab, msk = data
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(tiles))   # one shuffle; batches are contiguous slices of the shuffled set
    L = Ln.Learner(params)
    L.set_data(tiles[perm], q[perm], lab[perm], msk[perm])
    nb = len(tiles) // batch
    losses = []
    t0 = time.time()
    s = 0
    while s < steps:
        if budget and s == 20:
            steps = max(21, min(steps, int(budget * 20 / (time.time() - t0))))
            log("%d steps fit the budget of %.0f s" % (steps, budget))
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * s / steps))
        losses.append(L.step((s % nb) * batch, batch, cur))
        if s % 100 == 0 or s == steps - 1:
            log("step %d loss %.4f (%.0f s)" % (s, float(np.mean(losses[-100:])), time.time() - t0))
        s += 1
    p = L.params()
    L.close()
    return p, losses


def evaluate(params, n_img=12, seed=30000, qs=(60, 100, 140, 180), confs=(0.8, 0.9, 0.95), log=print):
    """BD-rate (PSNR-Y) of the pruned search at each confidence against the exhaustive one (rates summed over the
    images, PSNR averaged) and encode times (the network included)."""
    spec = image_spec(n_img, seed, 0.5, 2.0, qs=(120,))
    images = [(A.synth_image(s, W, H), W, H) for W, H, _, s in spec]
    res = {}
    for name, conf in [("exhaustive", None)] + [("pruned@%g" % c, c) for c in confs]:
        enc = A.GpEncoder().init(A.GpConfig(qindex=120))
        if conf is not None:
            Ln.set_shortlist(enc, params, conf)
        bits, psnr, secs = Ln.rd_curve(enc, images, qs)
        enc.close()
        res[name] = dict(bits=bits, psnr=psnr, seconds=round(secs, 4))
        log("eval %s: bits %s, PSNR-Y %s, %.3f s" % (name, bits, ["%.3f" % v for v in psnr], secs))
    ex = res["exhaustive"]
    for name, r in res.items():
        if name != "exhaustive":
            r["bd_rate_psnr_y"] = round(float(Ln.bd_rate(ex["bits"], ex["psnr"], r["bits"], r["psnr"])), 3)
            r["speedup"] = round(ex["seconds"] / r["seconds"], 3)
            log("%s: BD-rate %.3f %%, speed-up %.2fx" % (name, r["bd_rate_psnr_y"], r["speedup"]))
    mp = sum(W * H for W, H, _, _ in spec) * len(qs) / 1e6
    return dict(qindex=list(qs), images=n_img, megapixels=round(mp, 2), curves=res)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sbs", type=int, default=2000000)
    ap.add_argument("--steps", type=int, default=60000)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--train-budget", type=float, default=None, help="seconds of training at most")
    ap.add_argument("--rd-tau", type=float, default=None, help="cost-aware node labels at this relative threshold")
    ap.add_argument("--rd-tau-root", type=float, default=None, help="the threshold of the 64x64 node (default: --rd-tau)")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "shortlist_params.npy"))
    a = ap.parse_args()
    log = lambda m: print(m, flush=True)
    data = labelled_data(a.sbs, log=log, rd_tau=a.rd_tau, rd_tau_root=a.rd_tau_root)
    params, losses = train(Ln.init_params(1), data, a.steps, a.batch, a.lr, log=log, budget=a.train_budget)
    np.save(a.out, params)
    ev = evaluate(params, log=log)
    summary = dict(superblocks=int(len(data[0])), steps=len(losses), batch=a.batch, lr=a.lr, rd_tau=a.rd_tau, rd_tau_root=a.rd_tau_root,
                   loss_first=float(np.mean(losses[:50])), loss_last=float(np.mean(losses[-50:])), eval=ev)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
