"""Float64 NumPy restatement of the pruning network (av1gpu_amd/csrc/k_learn.hip) and of the encoder's decisions from
its logits: the checker of the network kernels' activations and gradients and of the pruned records
(tests/test_gpu_learn.py).  Test infrastructure only."""
import numpy as np

from av1gpu_amd.learn import CIN, FOUR0, KP, LEVEL_BASE, MODE0, N_SPLIT, OWN0, param_count, unpack


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def input_tensor(tiles, q):
    t = np.asarray(tiles, np.float32).reshape(-1, 64, 64)
    x0 = np.empty((len(t), 64, 64, 2), np.float32)
    x0[..., 0] = bf16(t * np.float32(1.0 / 255.0) - np.float32(0.5))
    x0[..., 1] = bf16(np.asarray(q, np.float32) / np.float32(255.0))[:, None, None]
    return x0


def _im2col(x, cin):
    """x [n, H, H, cin] -> [n, H/2 * H/2, 9 cin] for the 3x3 stride-2 pad-1 convolution (tap-major, then channel)."""
    n, H = x.shape[0], x.shape[1]
    Ho = H // 2
    xp = np.zeros((n, H + 2, H + 2, cin), x.dtype)
    xp[:, 1:H + 1, 1:H + 1] = x
    cols = [xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2, :] for ky in range(3) for kx in range(3)]
    return np.stack(cols, axis=3).reshape(n, Ho * Ho, 9 * cin)


def ref_conv(x, w, b, l):
    """One convolution + ReLU in float64 on BF16 inputs and weights; output rounded to BF16 as the kernel stores it."""
    K = 9 * CIN[l]
    col = _im2col(x.astype(np.float64), CIN[l])
    wt = bf16(w[:, :K]).astype(np.float64)
    z = col @ wt.T + b.astype(np.float64)
    H = x.shape[1] // 2
    return bf16(np.maximum(z, 0.0)).reshape(x.shape[0], H, H, 32)   # bf16 takes float32 first, then rounds


def ref_fc(a4, w5, b5):
    return a4.reshape(len(a4), -1).astype(np.float64) @ w5.astype(np.float64).T + b5


def ref_grads(params, acts, label, lmask):
    """Float64 gradient of the mean masked BCE, backpropagated through the given (kernel) activations
    acts = [x0, a1, a2, a3, a4] (NHWC float32 holding BF16 values)."""
    convs, w5, b5 = unpack(params)
    n = len(label)
    z = ref_fc(acts[4], w5, b5)
    s = 1.0 / (1.0 + np.exp(-z))
    dlog = lmask * (s - label) / n
    g = np.zeros(param_count(), np.float64)
    gconv, gw5, gb5 = unpack(g)
    a4 = acts[4].reshape(n, -1).astype(np.float64)
    gw5[:] = dlog.T @ a4
    gb5[:] = dlog.sum(0)
    d = (dlog @ w5.astype(np.float64)) * (a4 > 0)
    d = d.reshape(n, 4, 4, 32)
    for l in (3, 2, 1, 0):
        x = acts[l].astype(np.float64)
        K = 9 * CIN[l]
        col = _im2col(x, CIN[l])
        dz = d.reshape(n, -1, 32)
        gconv[l][0][:, :K] = np.einsum("npc,npk->ck", dz, col)
        gconv[l][1][:] = dz.sum((0, 1))
        if l > 0:
            wk = convs[l][0][:, :K].astype(np.float64)
            dcol = dz @ wk                                   # [n, P, 9 cin]
            H = x.shape[1]
            Ho = H // 2
            dx = np.zeros((n, H + 2, H + 2, CIN[l]))
            dcol = dcol.reshape(n, Ho, Ho, 9, CIN[l])
            for t in range(9):
                ky, kx = t // 3, t % 3
                dx[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2, :] += dcol[:, :, :, t, :]
            d = dx[:, 1:H + 1, 1:H + 1] * (x > 0)
    loss = float(np.sum(lmask * (np.maximum(z, 0) - z * label + np.log1p(np.exp(-np.abs(z))))) / n)
    return loss, g


def decide(logits, W, H, conf):
    """The decisions the encoder takes from the logits of an image's superblocks (raster order): expand [nSb, 85] bool
    (a node's children are analysed), own [nSb, 85] bool (its N x N / 2:1 shapes are), four [nSb, 85] bool (its 4-way
    shapes are) and the luma-mode mask [nSb] (bit m: mode m screened)."""
    z = np.asarray(logits, np.float64)
    MiCols, MiRows = 2 * ((W + 7) >> 3), 2 * ((H + 7) >> 3)
    FW, FH = 4 * MiCols, 4 * MiRows
    sbc = (MiCols + 15) >> 4
    idx = np.arange(len(z))
    sx, sy = (idx % sbc) * 64, (idx // sbc) * 64
    p = 1.0 / (1.0 + np.exp(-z))
    expand = np.zeros((len(z), N_SPLIT), bool)
    own = np.zeros((len(z), N_SPLIT), bool)
    four = np.zeros((len(z), N_SPLIT), bool)
    for lv in range(4):
        n, per = 64 >> lv, 1 << lv
        for ny in range(per):
            for nx in range(per):
                nd = LEVEL_BASE[lv] + ny * per + nx
                edge = (sx + nx * n + n > FW) | (sy + ny * n + n > FH)
                expand[:, nd] = edge | (p[:, nd] > 1.0 - conf)
                four[:, nd] = edge | (p[:, FOUR0 + nd] > 1.0 - conf)
                own[:, nd] = edge | (p[:, OWN0 + nd] > 1.0 - conf) | (~expand[:, nd] & ~four[:, nd])
    pm = p[:, MODE0:]
    order = np.argsort(-pm, axis=1, kind="stable")      # ties: the lower mode first
    top, rest = order[:, :8], order[:, 8:]
    none = np.prod(1.0 - np.take_along_axis(pm, rest, 1), axis=1)
    mask = np.full(len(z), 0x1FFF, np.int64)
    bits = np.sum(1 << top, axis=1)
    mask[none >= conf] = bits[none >= conf]
    return expand, own, four, mask
