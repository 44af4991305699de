"""Seeded synthetic RGB-D sequences for the odometry workloads.

The renderer itself is native (``csrc/synth.c``); this module owns the scene
handle, the camera trajectories and the frame containers.  It replaces the
reference's dataset readers (``src/inputs/FileReader.cpp:42-57``,
``RawLogReader.cpp:70-117``) for tests and benchmarks.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _native

_DP = ctypes.POINTER(ctypes.c_double)


def _lib():
    lib = _native.load("synth")
    if not getattr(lib, "_typed", False):
        lib.synth_scene_size.restype = ctypes.c_int
        lib.synth_scene_create.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
        lib.synth_render.argtypes = [
            ctypes.c_void_p, _DP, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
        ]
        lib._typed = True
    return lib


@dataclass
class Camera:
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float


# Camera models of the benchmark configurations.
CAM_320 = Camera(320, 240, 240.0, 240.0, 159.5, 119.5)
CAM_640 = Camera(640, 480, 481.2, 481.2, 319.5, 239.5)
CAM_848 = Camera(848, 480, 424.0, 424.0, 423.5, 239.5)


@dataclass
class Frame:
    """One rendered view.  ``depth``/``rgb`` are the noisy sensor images;
    ``vert4``/``norm4``/``rgba`` the noise-free model prediction."""
    pose: np.ndarray            # 4x4 float64 camera-to-world
    depth: np.ndarray | None    # (H, W) uint16, 1000 units/m
    rgb: np.ndarray | None      # (H, W, 3) uint8
    vert4: np.ndarray | None    # (H, W, 4) float32, camera frame
    norm4: np.ndarray | None    # (H, W, 4) float32, camera frame
    rgba: np.ndarray | None     # (H, W, 4) uint8


class Scene:
    """A procedural room: 6 planes + ``nbox`` axis-aligned boxes."""

    def __init__(self, seed: int, nbox: int = 20):
        lib = _lib()
        self._buf = ctypes.create_string_buffer(lib.synth_scene_size())
        lib.synth_scene_create(ctypes.c_uint64(seed), nbox, self._buf)
        self.seed = seed

    def render(self, pose: np.ndarray, cam: Camera, *, noise_mm: float = 1.0,
               dropout: float = 0.02, noise_seed: int = 0, sensor: bool = True,
               model: bool = True) -> Frame:
        lib = _lib()
        T = np.ascontiguousarray(pose, dtype=np.float64)
        H, W = cam.height, cam.width
        depth = np.empty((H, W), np.uint16) if sensor else None
        rgb = np.empty((H, W, 3), np.uint8) if sensor else None
        vert4 = np.empty((H, W, 4), np.float32) if model else None
        norm4 = np.empty((H, W, 4), np.float32) if model else None
        rgba = np.empty((H, W, 4), np.uint8) if model else None

        def ptr(a):
            return None if a is None else a.ctypes.data_as(ctypes.c_void_p)

        lib.synth_render(self._buf, T.ctypes.data_as(_DP), W, H, cam.fx, cam.fy, cam.cx, cam.cy,
                         ptr(depth), ptr(rgb), ptr(vert4), ptr(norm4), ptr(rgba),
                         float(noise_mm), float(dropout), ctypes.c_uint64(noise_seed))
        return Frame(T.copy(), depth, rgb, vert4, norm4, rgba)


# ---------------------------------------------------------------- surfel models

def encode_color(rgb: np.ndarray) -> np.ndarray:
    """color.glsl encodeColor of 8-bit colours: float(r << 16 | g << 8 | b)."""
    c = np.asarray(rgb, dtype=np.int64)
    return ((c[..., 0] << 16) | (c[..., 1] << 8) | c[..., 2]).astype(np.float32)


def surfels_from_frame(frame: Frame, cam: Camera, *, step: int = 1, conf: float = 10.0, init_time: int = 0,
                       last_time: int = 0, conf_jitter: float = 0.0, seed: int = 0) -> np.ndarray:
    """A surfel model, (n, 12) float32 in the layout {x, y, z, confidence,
    colour, 0, init_time, last_time, nx, ny, nz, radius}, from a rendered
    frame's noise-free maps: every ``step``-th valid pixel (z > 0) in each
    direction becomes a surfel at its global position with its global normal,
    the frame's colour and the radius of surfels.glsl getRadius
    ((z / meanFocal) * sqrt2 / |n_z|, at most twice (z / meanFocal) * sqrt2)."""
    v = frame.vert4[::step, ::step].reshape(-1, 4).astype(np.float64)
    n = frame.norm4[::step, ::step].reshape(-1, 4).astype(np.float64)
    col = encode_color(frame.rgba[::step, ::step].reshape(-1, 4)[:, :3])
    ok = (v[:, 2] > 0) & np.isfinite(v[:, 2]) & np.isfinite(n[:, 2])
    v, n, col = v[ok], n[ok], col[ok]
    R, t = frame.pose[:3, :3], frame.pose[:3, 3]
    mean_focal = ((1.0 / abs(1.0 / cam.fx)) + (1.0 / abs(1.0 / cam.fy))) / 2.0
    radius = (v[:, 2] / mean_focal) * 1.41421356237
    radius_n = np.minimum(2.0 * radius, radius / np.maximum(np.abs(n[:, 2]), 1e-6))
    out = np.zeros((v.shape[0], 12), np.float32)
    out[:, 0:3] = v[:, :3] @ R.T + t
    c = np.full(v.shape[0], conf, np.float64)
    if conf_jitter:
        c = c + np.random.default_rng(seed).uniform(-conf_jitter, conf_jitter, v.shape[0])
    out[:, 3] = c
    out[:, 4] = col
    out[:, 6] = init_time
    out[:, 7] = last_time
    out[:, 8:11] = n[:, :3] @ R.T
    out[:, 11] = radius_n
    return out


# ---------------------------------------------------------------- SE(3) helpers

def so3_exp(w: np.ndarray) -> np.ndarray:
    th = float(np.linalg.norm(w))
    if th < 1e-15:
        return np.eye(3)
    k = np.asarray(w, np.float64) / th
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * Kx + (1 - math.cos(th)) * (Kx @ Kx)


def so3_log(R: np.ndarray) -> np.ndarray:
    c = (np.trace(R) - 1.0) / 2.0
    th = math.acos(max(-1.0, min(1.0, c)))
    if th < 1e-12:
        return np.zeros(3)
    v = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    return v * (th / (2 * math.sin(th)))


def make_pose(R: np.ndarray, t: np.ndarray) -> np.ndarray:
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T


def base_pose() -> np.ndarray:
    """The camera's rest pose: at (0, 0, -1.5) looking along +z."""
    return make_pose(np.eye(3), np.array([0.0, 0.0, -1.5]))


def perturb(T: np.ndarray, rng: np.random.Generator, max_trans: float, max_rot_deg: float) -> np.ndarray:
    """T composed with a random rigid motion of at most ``max_trans`` metres and
    ``max_rot_deg`` degrees (in the camera frame)."""
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    dt = d * max_trans * rng.uniform(0.3, 1.0)
    dR = so3_exp(a * math.radians(max_rot_deg) * rng.uniform(0.3, 1.0))
    return T @ make_pose(dR, dt)


def pose_error(Ta: np.ndarray, Tb: np.ndarray) -> tuple[float, float]:
    """(translation error in metres, rotation error in degrees) of Ta vs Tb."""
    E = np.linalg.inv(Tb) @ Ta
    return float(np.linalg.norm(E[:3, 3])), float(np.degrees(np.linalg.norm(so3_log(E[:3, :3]))))


def spline_trajectory(n: int, seed: int, max_step_m: float = 0.01, max_step_deg: float = 0.5,
                      bursts: bool = False, max_excursion_m: float = 0.8) -> list[np.ndarray]:
    """A smooth seeded trajectory of ``n`` camera-to-world poses around the rest
    pose: sums of sinusoids per axis, scaled so the largest per-frame step is
    ``max_step_m`` / ``max_step_deg`` (reached to within 1 %).  ``bursts``
    modulates the speed with short fast segments (fast-motion stress).

    The camera stays within ``max_excursion_m`` of its start on every axis,
    so it stays well inside the room.  A long sequence would need a wider
    excursion at the requested step; instead its sinusoids run at a higher
    frequency (the path folds back sooner), which keeps the steps at the bound.
    """
    rng = np.random.default_rng(seed)
    k = np.arange(n, dtype=np.float64)
    speed = np.ones(n)
    if bursts:
        for _ in range(max(1, n // 150)):
            c = rng.uniform(0, n)
            speed += 6.0 * np.exp(-0.5 * ((k - c) / 6.0) ** 2)
    s = np.cumsum(speed)
    s = s / s[-1] * n
    # the sinusoids' parameters, drawn once: (axis, harmonic) ->
    # (frequency factor, position amplitude, phase, angle amplitude, phase)
    terms = []
    for ax in range(3):
        for m in range(1, 4):
            terms.append((ax, m, rng.uniform(0.8, 1.5), rng.uniform(0.2, 0.5) / m, rng.uniform(0, 2 * math.pi),
                          rng.uniform(0.1, 0.3) / m, rng.uniform(0, 2 * math.pi)))

    def curves(fscale: float):
        pos = np.zeros((n, 3))
        ang = np.zeros((n, 3))
        for ax, m, fq, pa, pp, aa, ap in terms:
            om = 2 * math.pi * m / max(n, 2) * fq * fscale
            pos[:, ax] += pa * np.sin(om * s + pp)
            ang[:, ax] += aa * np.sin(om * s + ap)
        pos -= pos[0]
        ang -= ang[0]
        ang[:, 2] *= 0.3  # less roll
        return pos, ang

    fscale = 1.0
    for _ in range(30):
        pos, ang = curves(fscale)
        dp = np.linalg.norm(np.diff(pos, axis=0), axis=1).max() if n > 1 else 1.0
        pos *= max_step_m / dp
        ext = np.abs(pos).max()
        if ext <= max_excursion_m:
            break
        fscale *= ext / max_excursion_m * 1.02
    Rs = [so3_exp(a) for a in ang]
    if n > 1:
        # scale to the bound; the step angle is not linear in the scale, so
        # rescale until it lands within 1 % below the bound
        for _ in range(30):
            dr = max(np.degrees(np.linalg.norm(so3_log(Rs[i].T @ Rs[i + 1]))) for i in range(n - 1))
            if 0.99 * max_step_deg <= dr <= max_step_deg:
                break
            ang *= max_step_deg / dr * (1.0 - 2e-3)
            Rs = [so3_exp(a) for a in ang]
    B = base_pose()
    return [B @ make_pose(Rs[i], pos[i]) for i in range(n)]
