"""Train the learned pruning network on the GPU and evaluate it against the exhaustive search.

  python -m av1gpu_amd.train_shortlist [--sbs 1000000] [--steps 6000] [--out av1gpu_amd/shortlist_params.npy]

1. Labels: seeded synthetic images (sizes log-uniform over 0.2-4 MP, aspect 0.5-2, qindex 80/120/160, seeds from
   20000) are encoded with the exhaustive search; every superblock's leaves give its 85 node labels and 13 mode shares
   (learn.sb_labels), its 64x64 luma tile is the input.
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


def labelled_data(n_sbs, seed=20000, log=print):
    """Tiles, qindex, labels and masks of about n_sbs superblocks of exhaustively encoded images."""
    enc = A.GpEncoder().init(A.GpConfig(qindex=120))
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
        for f in frames:
            f.close()
        for p, (W, H, q, _), o in zip(planes, spec, out):
            lab, msk = Ln.sb_labels(o.blocks, W, H)
            t = Ln.sb_tiles(p[0], W, H)
            tiles.append(t); qs.append(np.full(len(t), q, np.uint8)); labs.append(lab); msks.append(msk)
            total += len(t)
        log("data: %d superblocks from %d images (%.0f s)" % (total, k, time.time() - t0))
    enc.close()
    return np.concatenate(tiles), np.concatenate(qs), np.concatenate(labs), np.concatenate(msks)


def train(params, data, steps, batch, lr, log=print, seed=7):
    tiles, q, lab, msk = data
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(tiles))   # one shuffle; batches are contiguous slices of the shuffled set
    L = Ln.Learner(params)
    L.set_data(tiles[perm], q[perm], lab[perm], msk[perm])
    nb = len(tiles) // batch
    losses = []
    t0 = time.time()
    for s in range(steps):
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * s / steps))
        losses.append(L.step((s % nb) * batch, batch, cur))
        if s % 100 == 0 or s == steps - 1:
            log("step %d loss %.4f (%.0f s)" % (s, float(np.mean(losses[-100:])), time.time() - t0))
    p = L.params()
    L.close()
    return p, losses


def evaluate(params, n_img=12, seed=30000, qs=(60, 100, 140, 180), conf=0.8, log=print):
    """BD-rate (PSNR-Y) of the pruned search against the exhaustive one, per-image curves summed, and timings."""
    spec = image_spec(n_img, seed, 0.5, 2.0, qs=(120,))
    planes = [A.synth_image(s, W, H) for W, H, _, s in spec]
    res = {}
    for name, pr in (("exhaustive", None), ("pruned", params)):
        enc = A.GpEncoder().init(A.GpConfig(qindex=120))
        if pr is not None:
            Ln.set_shortlist(enc, pr, conf)
        rates, psnrs, secs, blocks = [], [], 0.0, 0
        for q in qs:
            frames = [A.DeviceFrame(p, W, H, qindex=q) for p, (W, H, _, _) in zip(planes, spec)]
            enc.encode(frames[:2])
            A.lib().av1mi_device_sync()
            t0 = time.perf_counter()
            out, t = enc.encode(frames)
            secs += time.perf_counter() - t0
            bits, mse_w = 0, []
            for p, (W, H, _, _), o in zip(planes, spec, out):
                bits += 8 * len(A.write_bitstream(o, avif=False))
                mse_w.append(Ln.psnr_y(p[0], o.recon[0], W, H))
                blocks += len(o.blocks)
            for f in frames:
                f.close()
            rates.append(bits)
            psnrs.append(float(np.mean(mse_w)))
            log("eval %s q%d: %d bits, PSNR-Y %.3f dB, %.3f s" % (name, q, bits, psnrs[-1], secs))
        enc.close()
        res[name] = dict(bits=rates, psnr=psnrs, seconds=secs, blocks=blocks)
    mp = sum(W * H for W, H, _, _ in spec) * len(qs) / 1e6
    bd = Ln.bd_rate(res["exhaustive"]["bits"], res["exhaustive"]["psnr"], res["pruned"]["bits"], res["pruned"]["psnr"])
    out = dict(bd_rate_psnr_y=round(float(bd), 3), qindex=list(qs), images=n_img, megapixels=round(mp, 2),
               exhaustive=res["exhaustive"], pruned=res["pruned"],
               speedup=round(res["exhaustive"]["seconds"] / res["pruned"]["seconds"], 3), confidence=conf)
    log("BD-rate %.3f %%, speed-up %.2fx" % (bd, out["speedup"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sbs", type=int, default=1000000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "shortlist_params.npy"))
    a = ap.parse_args()
    log = lambda m: print(m, flush=True)
    data = labelled_data(a.sbs, log=log)
    params, losses = train(Ln.init_params(1), data, a.steps, a.batch, a.lr, log=log)
    np.save(a.out, params)
    ev = evaluate(params, log=log)
    summary = dict(superblocks=int(len(data[0])), steps=a.steps, batch=a.batch, lr=a.lr,
                   loss_first=float(np.mean(losses[:50])), loss_last=float(np.mean(losses[-50:])), eval=ev)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
