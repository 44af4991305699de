"""Learned warm start: the BF16 MFMA network of ``csrc/cnn_kernels.hip``.

The network sees the model frame and the current frame (depth + intensity at
80x60, built from the tracker's level-2 pyramids or from 160x120 renders) and
predicts the camera motion between them as a scaled se(3) vector
``(t / T_SCALE, omega / R_SCALE)``.  The tracker starts its Gauss-Newton solve
from that motion (``RGBDOdometry.track_seeded``) instead of from the previous
pose, with a check that keeps the plain start when the prediction is worse.

No counterpart in the reference: its tracker always starts from the previous
pose (RGBDOdometryef.cpp:267-401), and its only network (YOLACT) is unrelated.
Training runs on the GPU with the library's own forward, backward and Adam
kernels; inputs come from the project's renderer, labels from its ground
truth motion.
"""
from __future__ import annotations

import ctypes
import math
import os
import time

import numpy as np

from . import _native, synth

T_SCALE = 0.05      # metres per output unit
R_SCALE = 0.05      # radians per output unit
IN_W, IN_H, IN_C = 80, 60, 4
CAM_160 = synth.Camera(160, 120, 120.3, 120.3, 79.5, 59.5)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart.npz")

_V = ctypes.c_void_p

# Every symbol include/slam_cnn.h declares.
EXPORTS = ["cnn_create", "cnn_destroy", "cnn_set_stream", "cnn_num_params", "cnn_param_layout", "cnn_set_params",
           "cnn_get_params", "cnn_get_grads", "cnn_prepare_input", "cnn_forward", "cnn_backward", "cnn_adam_step",
           "cnn_infer", "cnn_set_infer_split"]


def _lib():
    L = _native.load("odom")
    if not getattr(L, "_cnn_typed", False):
        for n in EXPORTS:
            getattr(L, n).restype = ctypes.c_int
        L.cnn_create.argtypes = [ctypes.c_int, ctypes.POINTER(_V)]
        L.cnn_destroy.argtypes = [_V]
        L.cnn_set_stream.argtypes = [_V, _V]
        L.cnn_num_params.argtypes = [_V, ctypes.POINTER(ctypes.c_size_t)]
        L.cnn_param_layout.argtypes = [_V, _V, _V, ctypes.c_int]
        L.cnn_set_params.argtypes = [_V, _V, ctypes.c_size_t]
        L.cnn_get_params.argtypes = [_V, _V, ctypes.c_size_t]
        L.cnn_get_grads.argtypes = [_V, _V, ctypes.c_size_t]
        L.cnn_prepare_input.argtypes = [_V, _V, _V, _V, _V, ctypes.c_int, ctypes.c_size_t, _V]
        L.cnn_forward.argtypes = [_V, _V, ctypes.c_int, _V]
        L.cnn_infer.argtypes = [_V, _V, ctypes.c_int, _V]
        L.cnn_set_infer_split.argtypes = [_V, ctypes.c_int]
        L.cnn_backward.argtypes = [_V, ctypes.c_int, _V, _V, _V]
        L.cnn_adam_step.argtypes = [_V, ctypes.c_float, ctypes.c_float]
        L._cnn_typed = True
    return L


class CnnError(RuntimeError):
    pass


def _check(fn: str, rc: int) -> int:
    if rc < 0:
        raise CnnError(f"{fn} returned {rc}")
    return rc


class WarmStartNet:
    """The network and its training state on the current GPU."""

    def __init__(self, max_batch: int = 256):
        L = _lib()
        h = _V()
        _check("cnn_create", L.cnn_create(int(max_batch), ctypes.byref(h)))
        self._h = h
        self.max_batch = max_batch
        n = ctypes.c_size_t()
        _check("cnn_num_params", L.cnn_num_params(h, ctypes.byref(n)))
        self.nparam = n.value
        off = (ctypes.c_size_t * 32)()
        sz = (ctypes.c_size_t * 32)()
        k = _check("cnn_param_layout", L.cnn_param_layout(h, off, sz, 32))
        self.layout = [(int(off[i]), int(sz[i])) for i in range(k)]

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr) -> None:
        _check("cnn_set_stream", _lib().cnn_set_stream(self._h, stream_ptr))

    def set_params(self, flat: np.ndarray) -> None:
        a = np.ascontiguousarray(flat, dtype=np.float32)
        if a.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {a.size}")
        _check("cnn_set_params", _lib().cnn_set_params(self._h, a.ctypes.data, a.size))

    def get_params(self) -> np.ndarray:
        a = np.zeros(self.nparam, np.float32)
        _check("cnn_get_params", _lib().cnn_get_params(self._h, a.ctypes.data, a.size))
        return a

    def get_grads(self) -> np.ndarray:
        a = np.zeros(self.nparam, np.float32)
        _check("cnn_get_grads", _lib().cnn_get_grads(self._h, a.ctypes.data, a.size))
        return a

    def init_params(self, seed: int = 0) -> None:
        """He-normal weights and zero biases."""
        self.set_params(he_init(self.layout, self.nparam, seed))

    def prepare(self, depth_a, image_a, depth_b, image_b):
        """Network input (bf16 [n][60][80][4]) from 160x120 device tensors."""
        import torch
        n = depth_a.shape[0]
        x = torch.empty((n, IN_H, IN_W, IN_C), dtype=torch.bfloat16, device=depth_a.device)
        _check("cnn_prepare_input", _lib().cnn_prepare_input(
            self._h, depth_a.data_ptr(), image_a.data_ptr(), depth_b.data_ptr(), image_b.data_ptr(), n,
            depth_a[0].numel(), x.data_ptr()))
        return x

    def forward(self, x, out=None):
        import torch
        n = x.shape[0]
        if out is None:
            out = torch.empty((n, 6), dtype=torch.float32, device=x.device)
        _check("cnn_forward", _lib().cnn_forward(self._h, x.data_ptr(), n, out.data_ptr()))
        return out

    def set_infer_split(self, split: bool) -> None:
        """Where infer computes the FC layers: in 16 workgroups per sample
        after the convolutions' launch (True, the default) or in the
        convolutions' workgroup (False); the same bits."""
        _check("cnn_set_infer_split", _lib().cnn_set_infer_split(self._h, int(bool(split))))

    def infer(self, x, out=None):
        """Inference (cnn_infer): the convolutions in one workgroup per sample,
        then the FC layers; no activations kept."""
        import torch
        n = x.shape[0]
        if out is None:
            out = torch.empty((n, 6), dtype=torch.float32, device=x.device)
        _check("cnn_infer", _lib().cnn_infer(self._h, x.data_ptr(), n, out.data_ptr()))
        return out

    def backward(self, n: int, target, weights=(1.0,) * 6):
        import torch
        w = np.ascontiguousarray(weights, dtype=np.float32)
        loss = torch.empty(n, dtype=torch.float32, device=target.device)
        _check("cnn_backward", _lib().cnn_backward(self._h, n, target.data_ptr(), w.ctypes.data, loss.data_ptr()))
        return loss

    def adam(self, lr: float, weight_decay: float = 0.0) -> None:
        _check("cnn_adam_step", _lib().cnn_adam_step(self._h, float(lr), float(weight_decay)))

    def save(self, path: str, meta: dict | None = None) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        m = dict(meta or {})
        np.savez_compressed(path, params=self.get_params(), t_scale=np.float64(T_SCALE),
                            r_scale=np.float64(R_SCALE),
                            meta=np.array(repr(sorted(m.items()))))

    def load(self, path: str = WEIGHTS) -> None:
        with np.load(path, allow_pickle=False) as z:
            p = z["params"]
            if float(z["t_scale"]) != T_SCALE or float(z["r_scale"]) != R_SCALE:
                raise ValueError("weights were trained with other output scales")
        self.set_params(p)


def he_init(layout, nparam: int, seed: int = 0) -> np.ndarray:
    """He-normal initial parameters in the library's flat layout (padding rows
    and the unused fc2 columns zero)."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(nparam, np.float32)
    cins = [4, 32, 64, 96, 128, 128]
    couts = [32, 64, 96, 128, 128, 128]
    for i in range(6):
        off, n = layout[2 * i]
        kp = n // couts[i]
        W = np.zeros((kp, couts[i]), np.float32)
        W[:9 * cins[i]] = rng.normal(0.0, math.sqrt(2.0 / (9 * cins[i])), size=(9 * cins[i], couts[i]))
        flat[off:off + n] = W.ravel()
    off, n = layout[12]
    flat[off:off + n] = rng.normal(0.0, math.sqrt(2.0 / 128), size=n)
    off, n = layout[14]
    hidden = n // 8
    W2 = np.zeros((hidden, 8), np.float32)
    W2[:, :6] = rng.normal(0.0, math.sqrt(1.0 / hidden), size=(hidden, 6))
    flat[off:off + n] = W2.ravel()
    return flat



# ----------------------------------------------------------------- data
def intensity(img: np.ndarray) -> np.ndarray:
    """The tracker's intensity of 3/4-byte texels (0.114, 0.299, 0.587 on bytes 0, 1, 2)."""
    f = img[..., :3].astype(np.float32)
    v = np.float32(0.114) * f[..., 0] + np.float32(0.299) * f[..., 1] + np.float32(0.587) * f[..., 2]
    return v.astype(np.uint8)


def motion_target(M: np.ndarray) -> np.ndarray:
    """Scaled se(3) label of a relative camera motion M (4x4): (t, rotvec)."""
    return np.concatenate([M[:3, 3] / T_SCALE, synth.so3_log(M[:3, :3]) / R_SCALE]).astype(np.float32)


def motion_from_pred(pred: np.ndarray) -> np.ndarray:
    p = np.asarray(pred, np.float64)
    return synth.make_pose(synth.so3_exp(p[3:6] * R_SCALE), p[:3] * T_SCALE)


def render_pairs(n: int, seed: int, max_trans: float = 0.08, max_rot_deg: float = 6.0, cam=CAM_160,
                 log=None) -> dict:
    """n seeded pairs: model frame A (noise-free depth and colour at pose T),
    current frame B (sensor depth and colour at T M), M random up to
    max_trans / max_rot_deg; T varies around the rest pose."""
    rng = np.random.default_rng(seed)
    H, W = cam.height, cam.width
    out = {"depth_a": np.zeros((n, H, W), np.float32), "image_a": np.zeros((n, H, W), np.uint8),
           "depth_b": np.zeros((n, H, W), np.uint16), "image_b": np.zeros((n, H, W), np.uint8),
           "target": np.zeros((n, 6), np.float32)}
    scene = None
    t0 = time.time()
    for i in range(n):
        if i % 16 == 0:
            scene = synth.Scene(int(rng.integers(1 << 30)))
        T = synth.perturb(synth.base_pose(), rng, 0.5, 20.0)
        M = synth.perturb(np.eye(4), rng, max_trans, max_rot_deg)
        fa = scene.render(T, cam, noise_seed=int(rng.integers(1 << 30)), sensor=False)
        fb = scene.render(T @ M, cam, noise_seed=int(rng.integers(1 << 30)), model=False)
        za = fa.vert4[..., 2]
        out["depth_a"][i] = np.where(za > 0, za, 0.0)
        out["image_a"][i] = intensity(fa.rgba)
        out["depth_b"][i] = fb.depth
        out["image_b"][i] = intensity(fb.rgb)
        out["target"][i] = motion_target(M)
        if log and (i + 1) % 2000 == 0:
            log(f"rendered {i + 1}/{n} pairs ({time.time() - t0:.0f}s)")
    return out


def _render_chunk(args):
    n, seed, max_trans, max_rot_deg = args
    return render_pairs(n, seed, max_trans, max_rot_deg)


def render_pairs_parallel(n: int, seed: int, max_trans: float = 0.08, max_rot_deg: float = 6.0,
                          workers: int = 16, chunk: int = 512, log=None) -> dict:
    """render_pairs in parallel worker processes (one renderer thread each),
    chunk c seeded with seed * 100003 + c."""
    import multiprocessing as mp
    import os as _os
    nchunks = (n + chunk - 1) // chunk
    jobs = [(min(chunk, n - c * chunk), seed * 100003 + c, max_trans, max_rot_deg) for c in range(nchunks)]
    old = _os.environ.get("OMP_NUM_THREADS")
    _os.environ["OMP_NUM_THREADS"] = "1"
    try:
        ctx = mp.get_context("spawn")
        with ctx.Pool(workers) as pool:
            parts = []
            t0 = time.time()
            for k, part in enumerate(pool.imap(_render_chunk, jobs)):
                parts.append(part)
                if log and (k + 1) % 16 == 0:
                    log(f"rendered {min(n, (k + 1) * chunk)}/{n} pairs ({time.time() - t0:.0f}s)")
    finally:
        if old is None:
            del _os.environ["OMP_NUM_THREADS"]
        else:
            _os.environ["OMP_NUM_THREADS"] = old
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def to_device(data: dict):
    import torch
    return {k: torch.from_numpy(v).cuda() for k, v in data.items()}


def train(net: WarmStartNet, train_dev: dict, val_dev: dict | None, epochs: int = 20, batch: int = 256,
          lr: float = 2e-3, weight_decay: float = 1e-4, seed: int = 0, log=None) -> list:
    """Adam on the weighted squared error of the scaled se(3) labels, cosine
    learning-rate decay; returns the per-epoch history."""
    import torch
    x_all = net.prepare(train_dev["depth_a"], train_dev["image_a"], train_dev["depth_b"], train_dev["image_b"])
    y_all = train_dev["target"]
    xv = yv = None
    if val_dev is not None:
        xv = net.prepare(val_dev["depth_a"], val_dev["image_a"], val_dev["depth_b"], val_dev["image_b"])
        yv = val_dev["target"]
    n = x_all.shape[0]
    steps = epochs * (n // batch)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).cuda()
        tot = 0.0
        for b in range(n // batch):
            idx = perm[b * batch:(b + 1) * batch]
            xb = x_all.index_select(0, idx).contiguous()
            yb = y_all.index_select(0, idx).contiguous()
            net.forward(xb)
            loss = net.backward(batch, yb)
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, steps)))
            net.adam(cur, weight_decay)
            tot += float(loss.mean())
            step += 1
        rec = {"epoch": ep, "train_loss": tot / max(1, n // batch)}
        if xv is not None:
            rec.update(evaluate(net, xv, yv, batch))
        hist.append(rec)
        if log:
            log(f"epoch {ep}: " + ", ".join(f"{k} {v:.4g}" for k, v in rec.items() if k != "epoch"))
    return hist


def evaluate(net: WarmStartNet, x, y, batch: int = 256) -> dict:
    """Validation loss and the mean translation / rotation error of the
    predicted motion against the labels (and of the zero-motion guess)."""
    import torch
    preds = []
    for b in range(0, x.shape[0], batch):
        xb = x[b:b + batch].contiguous()
        preds.append(net.forward(xb).clone())
    p = torch.cat(preds).cpu().numpy().astype(np.float64)
    t = y.cpu().numpy().astype(np.float64)
    et = np.linalg.norm((p[:, :3] - t[:, :3]) * T_SCALE, axis=1)
    er = np.degrees(np.linalg.norm((p[:, 3:] - t[:, 3:]) * R_SCALE, axis=1))
    zt = np.linalg.norm(t[:, :3] * T_SCALE, axis=1)
    zr = np.degrees(np.linalg.norm(t[:, 3:] * R_SCALE, axis=1))
    return {"val_loss": float(np.mean(np.sum((p - t) ** 2, axis=1))),
            "val_trans_err_mm": float(1e3 * np.mean(et)), "val_rot_err_deg": float(np.mean(er)),
            "zero_guess_trans_mm": float(1e3 * np.mean(zt)), "zero_guess_rot_deg": float(np.mean(zr))}
