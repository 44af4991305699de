"""FP64 numpy restatement of the warm-start network (slam_amd/csrc/cnn_kernels.hip).

TEST INFRASTRUCTURE ONLY: the checker of the HIP forward/backward kernels.
Only ``tests/`` may import it.

Same architecture and flat parameter layout as the library: six 3x3 / stride 2
/ pad 1 convolutions with ReLU (4-32-64-96-128-128-128 channels, NHWC), global
average pooling, FC 128-3840 (ReLU), FC 3840-6.  Per convolution W is [Kp][Cout]
with rows k = (kh * 3 + kw) * Cin + c (rows >= 9 Cin are padding), then b;
fc1 W [128][3840], b [3840]; fc2 W [3840][8], b [8] (columns 6, 7 unused).
Loss: sum_j w_j (out_j - target_j)^2 per sample, gradients of its mean.
"""
from __future__ import annotations

import numpy as np

SPEC = [(4, 32, 60, 80, 30, 40), (32, 64, 30, 40, 15, 20), (64, 96, 15, 20, 8, 10),
        (96, 128, 8, 10, 4, 5), (128, 128, 4, 5, 2, 3), (128, 128, 2, 3, 1, 2)]
FC1_IN, FC1_OUT, OUT, OUT_PAD = 128, 3840, 6, 8


def kp(cin: int) -> int:
    return ((9 * cin + 31) // 32) * 32


def layout():
    """(name, offset, shape) of every tensor in the flat vector (64-aligned offsets)."""
    out, off = [], 0

    def take(name, shape):
        nonlocal off
        n = int(np.prod(shape))
        out.append((name, off, shape))
        off += (n + 63) & ~63

    for i, (cin, cout, *_ ) in enumerate(SPEC):
        take(f"conv{i}.W", (kp(cin), cout))
        take(f"conv{i}.b", (cout,))
    take("fc1.W", (FC1_IN, FC1_OUT))
    take("fc1.b", (FC1_OUT,))
    take("fc2.W", (FC1_OUT, OUT_PAD))
    take("fc2.b", (OUT_PAD,))
    return out, off


def unpack(flat):
    lay, _ = layout()
    return {name: np.asarray(flat[o:o + int(np.prod(s))], np.float64).reshape(s) for name, o, s in lay}


def pack(tensors, total=None):
    lay, n = layout()
    flat = np.zeros(n if total is None else total, np.float64)
    for name, o, s in lay:
        flat[o:o + int(np.prod(s))] = np.asarray(tensors[name], np.float64).ravel()
    return flat


def im2col(x, ho, wo, kpad):
    """x: [N][H][W][C] -> [N*ho*wo][kpad], k = (kh*3+kw)*C + c."""
    n, h, w, c = x.shape
    xp = np.zeros((n, h + 2, w + 2, c), x.dtype)
    xp[:, 1:h + 1, 1:w + 1] = x
    cols = np.zeros((n, ho, wo, kpad), x.dtype)
    for kh in range(3):
        for kw in range(3):
            patch = xp[:, kh:kh + 2 * ho:2, kw:kw + 2 * wo:2, :]
            cols[..., (kh * 3 + kw) * c:(kh * 3 + kw + 1) * c] = patch[:, :ho, :wo]
    return cols.reshape(n * ho * wo, kpad)


def col2im(dcol, shape, ho, wo):
    n, h, w, c = shape
    d = dcol.reshape(n, ho, wo, -1)
    dxp = np.zeros((n, h + 2, w + 2, c), dcol.dtype)
    for kh in range(3):
        for kw in range(3):
            blk = d[..., (kh * 3 + kw) * c:(kh * 3 + kw + 1) * c]
            hh = min(ho, (h + 2 - kh + 1) // 2)
            ww = min(wo, (w + 2 - kw + 1) // 2)
            dxp[:, kh:kh + 2 * hh:2, kw:kw + 2 * ww:2, :] += blk[:, :hh, :ww]
    return dxp[:, 1:h + 1, 1:w + 1]


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return a


def forward(params, x, rnd=_ident):
    """params: dict from unpack(); x [N][60][80][4].  Returns (out [N][6], cache).
    rnd: applied where the library stores bf16 (each layer's output, the
    pooled features); identity for the exact network."""
    cache = {"x": [], "cols": [], "y": []}
    a = np.asarray(x, np.float64)
    for i, (cin, cout, hi, wi, ho, wo) in enumerate(SPEC):
        cols = im2col(a, ho, wo, kp(cin))
        y = rnd(np.maximum(cols @ params[f"conv{i}.W"] + params[f"conv{i}.b"], 0.0))
        cache["x"].append(a)
        cache["cols"].append(cols)
        a = y.reshape(a.shape[0], ho, wo, cout)
        cache["y"].append(a)
    g = rnd(a.reshape(a.shape[0], -1, FC1_IN).mean(axis=1))
    h1 = rnd(np.maximum(g @ params["fc1.W"] + params["fc1.b"], 0.0))
    out = h1 @ params["fc2.W"] + params["fc2.b"]
    cache.update(g=g, h1=h1, out=out)
    return out[:, :OUT], cache


def loss_and_grads(params, x, target, w, rnd=_ident):
    """Mean over samples of sum_j w_j (out_j - t_j)^2 and its gradients (dict).
    rnd as in forward, also applied to the gradients the library feeds its
    GEMMs in bf16 (bias gradients use the unrounded values, as there)."""
    out, c = forward(params, x, rnd)
    n = out.shape[0]
    e = out - np.asarray(target, np.float64)
    w = np.asarray(w, np.float64)
    loss = float(np.mean(np.sum(w * e * e, axis=1)))
    dout = np.zeros((n, OUT_PAD))
    dout[:, :OUT] = 2.0 * w * e / n
    g = {}
    g["fc2.W"] = c["h1"].T @ rnd(dout)
    g["fc2.b"] = dout.sum(0)
    dh1 = (rnd(dout) @ params["fc2.W"].T) * (c["h1"] > 0)
    g["fc1.W"] = c["g"].T @ rnd(dh1)
    g["fc1.b"] = dh1.sum(0)
    dg = rnd(dh1) @ params["fc1.W"].T
    last = c["y"][-1]
    p = last.shape[1] * last.shape[2]
    dy = np.repeat(dg[:, None, :], p, axis=1).reshape(last.shape) / p
    for i in range(len(SPEC) - 1, -1, -1):
        cin, cout, hi, wi, ho, wo = SPEC[i]
        dz = (dy * (c["y"][i] > 0)).reshape(-1, cout)
        g[f"conv{i}.W"] = c["cols"][i].T @ rnd(dz)
        g[f"conv{i}.b"] = dz.sum(0)
        if i == 0:
            break
        dcol = rnd(dz) @ params[f"conv{i}.W"].T
        dy = col2im(dcol, c["x"][i].shape, ho, wo)
    # padding entries carry no gradient
    for i, (cin, *_ ) in enumerate(SPEC):
        g[f"conv{i}.W"][9 * cin:] = 0.0
    g["fc2.W"][:, OUT:] = 0.0
    g["fc2.b"][OUT:] = 0.0
    return loss, g


def init_params(seed: int = 0):
    """He-normal weights, zero biases, zero padding rows."""
    rng = np.random.default_rng(seed)
    t = {}
    for i, (cin, cout, *_ ) in enumerate(SPEC):
        W = np.zeros((kp(cin), cout))
        W[:9 * cin] = rng.normal(0.0, np.sqrt(2.0 / (9 * cin)), size=(9 * cin, cout))
        t[f"conv{i}.W"] = W
        t[f"conv{i}.b"] = np.zeros(cout)
    t["fc1.W"] = rng.normal(0.0, np.sqrt(2.0 / FC1_IN), size=(FC1_IN, FC1_OUT))
    t["fc1.b"] = np.zeros(FC1_OUT)
    W2 = np.zeros((FC1_OUT, OUT_PAD))
    W2[:, :OUT] = rng.normal(0.0, np.sqrt(1.0 / FC1_OUT), size=(FC1_OUT, OUT))
    t["fc2.W"] = W2
    t["fc2.b"] = np.zeros(OUT_PAD)
    return t
