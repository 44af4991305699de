"""Surfel splat prediction and fill-in over ``libslam_odom.so``
(``include/slam_splat.h``): the producers of the model maps the odometry
consumes.

``IndexMap`` keeps the reference's names: ``combinedPredict`` is
``IndexMap::combinedPredict`` (``IndexMap.cpp:243-345``) and ``fillIn`` is
``FillIn::vertex`` / ``normal`` / ``image`` (``FillIn.cpp:56-198``) on the
last prediction.  Outputs are read back as numpy arrays or handed to the
odometry as device images (``DeviceImage``) without a copy.  There is no
Python or CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import odometry
from .odometry import DeviceImage, FMT_DEPTH_U16, MEM_DEVICE, MEM_HOST, _check, _p, _rgb_fmt, frame_view

# Every symbol include/slam_splat.h declares.
EXPORTS = ["splat_create", "splat_destroy", "splat_set_stream", "splat_combined_predict", "splat_fill_in",
           "splat_get_buffer", "splat_read", "splat_synchronize", "splat_profile"]

IMAGE, VERTEX, NORMAL, TIME, FILL_IMAGE, FILL_VERTEX, FILL_NORMAL = range(7)
SURFEL_FLOATS = 12
MAX_POINT_SIZE = 256

_V = ctypes.c_void_p
_P = ctypes.POINTER
_SHAPE = {IMAGE: (4, np.uint8), VERTEX: (4, np.float32), NORMAL: (4, np.float32), TIME: (1, np.uint16),
          FILL_IMAGE: (4, np.uint8), FILL_VERTEX: (4, np.float32), FILL_NORMAL: (4, np.float32)}


def lib() -> ctypes.CDLL:
    L = odometry.lib()
    if getattr(L, "_splat_typed", False):
        return L
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    F, I = ctypes.c_float, ctypes.c_int
    L.splat_create.argtypes = [I, I, F, F, F, F, I, _P(_V)]
    L.splat_destroy.argtypes = [_V]
    L.splat_set_stream.argtypes = [_V, _V]
    L.splat_combined_predict.argtypes = [_V, _V, _V, I, I, F, F, I, I, I]
    L.splat_fill_in.argtypes = [_V, _P(odometry.FrameView), _P(odometry.FrameView), I, I]
    L.splat_get_buffer.argtypes = [_V, I, _P(_V)]
    L.splat_read.argtypes = [_V, I, _V]
    L.splat_synchronize.argtypes = [_V]
    L.splat_profile.argtypes = [_V, _V, _V, I, F, F, I, I, I, _P(odometry.FrameView), _P(odometry.FrameView), I,
                                _P(F), _P(ctypes.c_longlong)]
    L._splat_typed = True
    return L


def _model_arg(model):
    """(pointer, count, memory kind, keep-alive) of an (n, 12) float32 model:
    a numpy array (host), a torch CUDA tensor, or a (device pointer, count)
    pair such as ``SurfelModel.device_model()``."""
    if isinstance(model, tuple):
        ptr, n = model
        return int(ptr), int(n), MEM_DEVICE, None
    try:
        import torch
        if isinstance(model, torch.Tensor) and model.is_cuda:
            t = model.contiguous()
            if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != SURFEL_FLOATS:
                raise ValueError("model must be (n, 12) float32")
            return t.data_ptr(), t.shape[0], MEM_DEVICE, t
    except ImportError:  # pragma: no cover
        pass
    a = np.ascontiguousarray(model, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != SURFEL_FLOATS:
        raise ValueError("model must be (n, 12) float32")
    return a.ctypes.data, a.shape[0], MEM_HOST, a


class IndexMap:
    """One prediction target (the frame loop's ACTIVE or INACTIVE buffers)
    with its fill-in textures."""

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 max_host_surfels: int = 0):
        self.width, self.height = width, height
        h = _V()
        _check("splat_create", lib().splat_create(width, height, cx, cy, fx, fy, max_host_surfels, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().splat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass

    def set_stream(self, stream: int | None) -> None:
        _check("splat_set_stream", lib().splat_set_stream(self._h, stream or None))

    def combinedPredict(self, pose, model, depthCutoff: float, confThreshold: float, time: int, maxTime: int,
                        timeDelta: int) -> None:
        T = np.ascontiguousarray(pose, dtype=np.float32).reshape(16)
        ptr, n, mem, keep = _model_arg(model)
        _check("splat_combined_predict",
               lib().splat_combined_predict(self._h, _p(T), ptr if n else None, n, mem, depthCutoff, confThreshold,
                                            time, maxTime, timeDelta))
        self._keep_model = keep

    def fillIn(self, depth, rgb, passthroughGeometry: bool = False, passthroughImage: bool = False) -> None:
        d, kd = frame_view(depth, FMT_DEPTH_U16)
        r, kr = frame_view(rgb, _rgb_fmt(rgb))
        _check("splat_fill_in", lib().splat_fill_in(self._h, ctypes.byref(d), ctypes.byref(r),
                                                    int(passthroughGeometry), int(passthroughImage)))
        self._keep_fill = (kd, kr)

    def read(self, which: int) -> np.ndarray:
        ch, dt = _SHAPE[which]
        shape = (self.height, self.width) if ch == 1 else (self.height, self.width, ch)
        out = np.empty(shape, dt)
        _check("splat_read", lib().splat_read(self._h, which, _p(out)))
        return out

    def device(self, which: int) -> DeviceImage:
        ptr = _V()
        _check("splat_get_buffer", lib().splat_get_buffer(self._h, which, ctypes.byref(ptr)))
        return DeviceImage(ptr.value, self.width, self.height, owner=self)

    # the reference's accessors, as device images for the odometry's init calls
    def imageTex(self) -> DeviceImage:
        return self.device(IMAGE)

    def vertexTex(self) -> DeviceImage:
        return self.device(VERTEX)

    def normalTex(self) -> DeviceImage:
        return self.device(NORMAL)

    def synchronize(self) -> None:
        _check("splat_synchronize", lib().splat_synchronize(self._h))

    def profile(self, pose, model_dev, depthCutoff, confThreshold, time, maxTime, timeDelta, depth=None, rgb=None,
                reps: int = 20) -> dict:
        """Per-launch microseconds of the raster, resolve and fill-in passes
        (HIP events on the handle's stream) and the fragments drawn."""
        T = np.ascontiguousarray(pose, dtype=np.float32).reshape(16)
        ptr, n, mem, keep = _model_arg(model_dev)
        if mem != MEM_DEVICE:
            raise ValueError("profile takes a device model")
        dv = rv = None
        if depth is not None:
            dv, kd = frame_view(depth, FMT_DEPTH_U16)
            rv, kr = frame_view(rgb, _rgb_fmt(rgb))
        us = (ctypes.c_float * 3)()
        frags = ctypes.c_longlong(0)
        _check("splat_profile", lib().splat_profile(self._h, _p(T), ptr, n, depthCutoff, confThreshold, time, maxTime,
                                                    timeDelta, ctypes.byref(dv) if dv else None,
                                                    ctypes.byref(rv) if rv else None, reps, us, ctypes.byref(frags)))
        return {"raster_us": us[0], "resolve_us": us[1], "fill_us": us[2], "fragments": int(frags.value)}
