"""The surfel model's fusion over ``libslam_odom.so`` (``include/slam_surfel.h``).

``SurfelModel`` keeps the reference's call names: ``initialise``
(``Model::initialise`` from the first frame's feedback buffers),
``predictIndices`` (``IndexMap::predictIndices``), ``fuse``
(``Model::fuse``), ``clean`` (``Model::clean`` without deformation nodes),
``lastCount``.  Frames are numpy arrays (host) or torch CUDA tensors
(device); the model stays on the device and feeds the splat prediction
directly (``device_model``).  There is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import odometry
from .odometry import FMT_DEPTH_U16, FMT_RGB8, MEM_DEVICE, MEM_HOST, _check, _p, frame_view

# Every symbol include/slam_surfel.h declares.
EXPORTS = ["surfel_create", "surfel_destroy", "surfel_set_stream", "surfel_initialise", "surfel_predict_indices",
           "surfel_fuse", "surfel_clean", "surfel_count", "surfel_get_model", "surfel_read_model",
           "surfel_set_model", "surfel_read_index_map", "surfel_read_fresh", "surfel_profile"]

_V = ctypes.c_void_p
_P = ctypes.POINTER


def lib() -> ctypes.CDLL:
    L = odometry.lib()
    if getattr(L, "_surfel_typed", False):
        return L
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    F, I, FV = ctypes.c_float, ctypes.c_int, _P(odometry.FrameView)
    L.surfel_create.argtypes = [I, I, F, F, F, F, I, _P(_V)]
    L.surfel_destroy.argtypes = [_V]
    L.surfel_set_stream.argtypes = [_V, _V]
    L.surfel_initialise.argtypes = [_V, FV, FV, FV, F, I, F]
    L.surfel_predict_indices.argtypes = [_V, _V, I, F, I]
    L.surfel_fuse.argtypes = [_V, _V, I, FV, FV, FV, F, F, F]
    L.surfel_clean.argtypes = [_V, _V, I, F, I]
    L.surfel_count.argtypes = [_V, _P(I)]
    L.surfel_get_model.argtypes = [_V, _P(_V)]
    L.surfel_read_model.argtypes = [_V, _V, I, _P(I)]
    L.surfel_set_model.argtypes = [_V, _V, I, I]
    L.surfel_read_index_map.argtypes = [_V, _V, _V, _V, _V]
    L.surfel_read_fresh.argtypes = [_V, _V, I, _P(I)]
    L.surfel_profile.argtypes = [_V, _V, I, FV, FV, FV, F, F, F, F, I, I, _P(F)]
    L._surfel_typed = True
    return L


def _pose(T):
    return np.ascontiguousarray(T, dtype=np.float32).reshape(16)


class SurfelModel:
    """The global surfel model (two buffers of ``max_surfels``), its index
    map and the fusion scratch, on one device."""

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float, max_surfels: int):
        self.width, self.height, self.max_surfels = width, height, max_surfels
        h = _V()
        _check("surfel_create", lib().surfel_create(width, height, cx, cy, fx, fy, max_surfels, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().surfel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass

    def set_stream(self, stream: int | None) -> None:
        _check("surfel_set_stream", lib().surfel_set_stream(self._h, stream or None))

    def _frames(self, rgb, raw, filtered):
        r, kr = frame_view(rgb, FMT_RGB8)
        a, ka = frame_view(raw, FMT_DEPTH_U16)
        f, kf = frame_view(filtered, FMT_DEPTH_U16)
        self._keep = (kr, ka, kf)
        return ctypes.byref(r), ctypes.byref(a), ctypes.byref(f)

    def initialise(self, rgb, rawDepth, filteredDepth, depthCutoff: float, time: int, maxDepth: float) -> int:
        _check("surfel_initialise", lib().surfel_initialise(self._h, *self._frames(rgb, rawDepth, filteredDepth),
                                                            depthCutoff, time, maxDepth))
        return self.lastCount()

    def predictIndices(self, pose, time: int, maxDepth: float, timeDelta: int) -> None:
        _check("surfel_predict_indices", lib().surfel_predict_indices(self._h, _p(_pose(pose)), time, maxDepth,
                                                                      timeDelta))

    def fuse(self, pose, time: int, rgb, rawDepth, filteredDepth, depthCutoff: float, maxDepth: float,
             weighting: float) -> None:
        _check("surfel_fuse", lib().surfel_fuse(self._h, _p(_pose(pose)), time,
                                                *self._frames(rgb, rawDepth, filteredDepth), depthCutoff, maxDepth,
                                                weighting))

    def clean(self, pose, time: int, confThreshold: float, timeDelta: int) -> int:
        _check("surfel_clean", lib().surfel_clean(self._h, _p(_pose(pose)), time, confThreshold, timeDelta))
        return self.lastCount()

    def lastCount(self) -> int:
        n = ctypes.c_int(0)
        _check("surfel_count", lib().surfel_count(self._h, ctypes.byref(n)))
        return n.value

    def device_model(self):
        """(device pointer, count) of the current model buffer."""
        p = _V()
        _check("surfel_get_model", lib().surfel_get_model(self._h, ctypes.byref(p)))
        return p.value, self.lastCount()

    def model(self) -> np.ndarray:
        n = self.lastCount()
        out = np.zeros((max(n, 1), 12), np.float32)
        got = ctypes.c_int(0)
        _check("surfel_read_model", lib().surfel_read_model(self._h, _p(out), n, ctypes.byref(got)))
        return out[:n]

    def set_model(self, model) -> None:
        try:
            import torch
            if isinstance(model, torch.Tensor) and model.is_cuda:
                t = model.contiguous()
                _check("surfel_set_model", lib().surfel_set_model(self._h, t.data_ptr(), t.shape[0], MEM_DEVICE))
                return
        except ImportError:  # pragma: no cover
            pass
        m = np.ascontiguousarray(model, dtype=np.float32).reshape(-1, 12)
        _check("surfel_set_model", lib().surfel_set_model(self._h, _p(m) if m.size else None, m.shape[0], MEM_HOST))

    def index_map(self):
        H, W = self.height, self.width
        idx = np.zeros((H, W), np.uint32)
        vc, ct, nr = (np.zeros((H, W, 4), np.float32) for _ in range(3))
        _check("surfel_read_index_map", lib().surfel_read_index_map(self._h, _p(idx), _p(vc), _p(ct), _p(nr)))
        return idx, vc, ct, nr

    def fresh(self) -> np.ndarray:
        cap = self.width * self.height
        out = np.zeros((cap, 12), np.float32)
        n = ctypes.c_int(0)
        _check("surfel_read_fresh", lib().surfel_read_fresh(self._h, _p(out), cap, ctypes.byref(n)))
        return out[:n.value]

    def profile(self, pose, time, rgb, rawDepth, filteredDepth, depthCutoff, maxDepth, weighting, confThreshold,
                timeDelta, reps: int = 20) -> dict:
        us = (ctypes.c_float * 4)()
        _check("surfel_profile", lib().surfel_profile(self._h, _p(_pose(pose)), time,
                                                      *self._frames(rgb, rawDepth, filteredDepth), depthCutoff,
                                                      maxDepth, weighting, confThreshold, timeDelta, reps, us))
        return {"predict_indices_us": us[0], "fuse_us": us[1], "predict_indices2_us": us[2], "clean_us": us[3]}
