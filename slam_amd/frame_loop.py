"""The reference's frame loop without loop closure, on the device: bilateral
filter, frame-to-model tracking against the filled-in splat prediction, the
model update, and the next prediction (elastic_fusion_file.cpp:335-420,
551-602, with closeLoops off).

``FrameLoop`` drives the native loop of ``include/slam_loop.h``
(``efloop_process``, C++ over the odometry, splat and surfel entry points);
with the learned warm start it runs the same steps from Python, adding the
network call.  Every stage runs on one HIP stream (the odometry's); the host
reads back only what the reference's loop reads: the tracked pose and the
model's surfel count.  The prediction that the reference makes right after
tracking (elastic_fusion_file.cpp:399) feeds only the loop-closure search, so
it is not made; the one at the end of the frame (:601) is.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import splat, surfels
from .odometry import (FMT_DEPTH_U16, FMT_RGB8, MEM_DEVICE, RGBDOdometry, Stats, _check, _p, frame_view,
                       lib as odom_lib)

# Every symbol include/slam_loop.h declares.
EXPORTS = ["efloop_default_params", "efloop_create", "efloop_destroy", "efloop_process", "efloop_get_handles",
           "efloop_get_state"]


class _Params(ctypes.Structure):
    _fields_ = [("depth_cutoff", ctypes.c_float), ("max_depth", ctypes.c_float), ("confidence", ctypes.c_float),
                ("time_delta", ctypes.c_int), ("icp_weight", ctypes.c_float), ("so3", ctypes.c_int),
                ("max_surfels", ctypes.c_int), ("use_graph", ctypes.c_int), ("autotune", ctypes.c_int)]


def lib() -> ctypes.CDLL:
    L = surfels.lib()
    splat.lib()
    if getattr(L, "_loop_typed", False):
        return L
    V, P = ctypes.c_void_p, ctypes.POINTER
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    L.efloop_default_params.argtypes = [P(_Params)]
    L.efloop_create.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 4 + [P(_Params), P(V)]
    L.efloop_destroy.argtypes = [V]
    L.efloop_process.argtypes = [V, P(odom_frame_view_type()), P(odom_frame_view_type()), V]
    L.efloop_get_handles.argtypes = [V, P(V), P(V), P(V)]
    L.efloop_get_state.argtypes = [V, P(ctypes.c_int), P(ctypes.c_float), P(ctypes.c_float)]
    L._loop_typed = True
    return L


def odom_frame_view_type():
    from .odometry import FrameView
    return FrameView


class _OdoView:
    """The loop's odometry handle (owned by the loop): statistics."""

    def __init__(self, h):
        self._h = h

    def stats(self) -> dict:
        s = Stats()
        _check("odom_get_stats", odom_lib().odom_get_stats(self._h, ctypes.byref(s)))
        return {"lastICPError": s.last_icp_error, "lastICPCount": s.last_icp_count,
                "lastRGBError": s.last_rgb_error, "lastRGBCount": s.last_rgb_count,
                "seedCalls": s.seed_calls, "seedKept": s.seed_kept}


class _ModelView:
    """The loop's surfel model (owned by the loop): count and a host copy."""

    def __init__(self, h):
        self._h = h

    def lastCount(self) -> int:
        n = ctypes.c_int(0)
        _check("surfel_count", surfels.lib().surfel_count(self._h, ctypes.byref(n)))
        return n.value

    def model(self) -> np.ndarray:
        n = self.lastCount()
        out = np.zeros((max(n, 1), 12), np.float32)
        got = ctypes.c_int(0)
        _check("surfel_read_model", surfels.lib().surfel_read_model(self._h, _p(out), n, ctypes.byref(got)))
        return out[:n]


@dataclass
class LoopParams:
    """The reference configs' values (src/configs/ef_file.cfg)."""
    depth_cutoff: float = 3.0        # depthCutoff: bilateral and metric depth window (m)
    max_depth: float = 20.0          # maxDepthProcessed
    confidence: float = 10.0
    time_delta: int = 200
    icp_weight: float = 10.0
    so3: bool = True
    max_surfels: int = 4_000_000
    use_graph: bool = True           # replay the tracking call from a hipGraph (native loop)
    autotune: bool = True            # step-kernel launch shapes timed on the first tracked frame
    learned: bool = False            # seed each call from the warm-start network (no SO3 pass then)
    learned_iters: tuple = (10, 5, 4)


def _rodrigues_norm(R: np.ndarray) -> float:
    c = (np.trace(R) - 1.0) / 2.0
    return float(np.arccos(np.clip(c, -1.0, 1.0)))


class FrameLoop:
    """Processes frames (uint16 depth in mm, RGB8), host or device."""

    def __new__(cls, width, height, cx, cy, fx, fy, params: LoopParams | None = None):
        if params is not None and params.learned:
            return super().__new__(_PythonFrameLoop)
        return super().__new__(cls)

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 params: LoopParams | None = None):
        self.p = params or LoopParams()
        P = _Params()
        _check("efloop_default_params", lib().efloop_default_params(ctypes.byref(P)))
        P.depth_cutoff, P.max_depth, P.confidence = self.p.depth_cutoff, self.p.max_depth, self.p.confidence
        P.time_delta, P.icp_weight, P.so3 = self.p.time_delta, self.p.icp_weight, int(self.p.so3)
        P.max_surfels = self.p.max_surfels
        P.use_graph = int(self.p.use_graph)
        P.autotune = int(self.p.autotune)
        h = ctypes.c_void_p()
        _check("efloop_create", lib().efloop_create(width, height, cx, cy, fx, fy, ctypes.byref(P), ctypes.byref(h)))
        self._h = h
        o, m = ctypes.c_void_p(), ctypes.c_void_p()
        _check("efloop_get_handles", lib().efloop_get_handles(h, ctypes.byref(o), None, ctypes.byref(m)))
        self.odo = _OdoView(o)
        self.model = _ModelView(m)
        self.poses: list[np.ndarray] = []

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().efloop_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass

    def process(self, rgb, depth) -> np.ndarray:
        """One frame; returns the tracked camera-to-model pose (4x4)."""
        r, kr = frame_view(rgb, FMT_RGB8)
        d, kd = frame_view(depth, FMT_DEPTH_U16)
        pose = np.zeros(16, np.float32)
        _check("efloop_process", lib().efloop_process(self._h, ctypes.byref(r), ctypes.byref(d), _p(pose)))
        T = pose.reshape(4, 4)
        self.poses.append(T.copy())
        return T


class _PythonFrameLoop(FrameLoop):
    """The same loop driven from Python, with the learned warm start seeding
    each tracking call (no SO3 pass then)."""

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 params: LoopParams | None = None):
        import torch
        self.p = params or LoopParams()
        self.w, self.h = width, height
        self.odo = RGBDOdometry(width, height, cx, cy, fx, fy)
        stream = self.odo.stream
        self.pred = splat.IndexMap(width, height, cx, cy, fx, fy)
        self.pred.set_stream(stream)
        self.model = surfels.SurfelModel(width, height, cx, cy, fx, fy, self.p.max_surfels)
        self.model.set_stream(stream)
        self.filtered = torch.empty((height, width), dtype=torch.uint16, device="cuda")
        self.pose = np.eye(4, dtype=np.float32)
        self.odo.set_pose(self.pose[:3, 3], self.pose[:3, :3])
        self.net = None
        self._h = None
        if self.p.learned:
            from . import warmstart as ws
            self.ws = ws
            self.net = ws.WarmStartNet(max_batch=1)
            self.net.load()
            self.net.set_stream(stream)
            self.x = torch.zeros((1, 60, 80, 4), dtype=torch.bfloat16, device="cuda")
            self.pred6 = torch.zeros((1, 6), dtype=torch.float32, device="cuda")
        self.tick = 0
        self.last_icp_error = 0.0
        self.poses: list[np.ndarray] = []

    def _bilateral(self, depth):
        v, keep = frame_view(depth, FMT_DEPTH_U16)
        _check("odom_bilateral_depth", odom_lib().odom_bilateral_depth(
            self.odo.handle, ctypes.byref(v), self.p.depth_cutoff, ctypes.c_void_p(self.filtered.data_ptr()),
            MEM_DEVICE))
        return self.filtered

    def _predict(self):
        """predict() (elastic_fusion_file.cpp:17-44): the active splat at the
        current pose, then the fill-in from the filtered depth and colour."""
        self.pred.combinedPredict(self.pose, self.model.device_model(), self.p.max_depth, self.p.confidence,
                                  self.tick, self.tick, self.p.time_delta)

    def process(self, rgb, depth) -> np.ndarray:
        """One frame; returns the tracked camera-to-model pose (4x4)."""
        self.tick += 1
        p = self.p
        filtered = self._bilateral(depth)
        if self.tick == 1:
            self.model.initialise(rgb, depth, filtered, p.depth_cutoff, self.tick, p.max_depth)
            self.odo.initFirstRGB(rgb)
        else:
            last = self.pose.copy()
            tracking_ok = self.last_icp_error < 1e-4      # the previous call's error, as the loop reads it
            self.odo.initICPModel(self.pred.device(splat.FILL_VERTEX), self.pred.device(splat.FILL_NORMAL),
                                  p.max_depth, self.pose)
            self.odo.initRGBModel(self.pred.device(splat.FILL_IMAGE))
            self.odo.initICP(filtered, p.max_depth)
            self.odo.initRGB(rgb)
            if self.net is not None:
                # the learned warm start: the network's motion seeds the call,
                # which keeps it only if its association is at least as good
                self.odo.initWarmStartInput(self.x)
                self.ws._lib().cnn_forward(self.net.handle, self.x.data_ptr(), 1, self.pred6.data_ptr())
                self.odo.track_seeded(self.pred6, p.learned_iters, False, p.icp_weight, True, self.ws.T_SCALE,
                                      self.ws.R_SCALE)
                t, R = self.odo.get_pose()
            else:
                t, R = self.odo.getIncrementalTransformation(self.pose[:3, 3], self.pose[:3, :3], False,
                                                             p.icp_weight, True, False, p.so3)
            self.pose = np.eye(4, dtype=np.float32)
            self.pose[:3, :3] = np.asarray(R, np.float32).reshape(3, 3)
            self.pose[:3, 3] = np.asarray(t, np.float32).reshape(3)
            self.last_icp_error = self.odo.stats()["lastICPError"]
            if self.tick == 2 and p.autotune:
                self.odo.autotune(False, p.icp_weight)  # launch shapes from the first tracked frame
            # weight by velocity (elastic_fusion_file.cpp:377-398)
            diff = np.linalg.inv(self.pose.astype(np.float64)) @ last.astype(np.float64)
            weighting = max(float(np.linalg.norm(diff[:3, 3])), _rodrigues_norm(diff[:3, :3]))
            weighting = max(1.0 - min(weighting, 0.01) / 0.01, 0.5)
            if tracking_ok:
                self.model.predictIndices(self.pose, self.tick, p.max_depth, p.time_delta)
                self.model.fuse(self.pose, self.tick, rgb, depth, filtered, p.depth_cutoff, p.max_depth, weighting)
                self.model.predictIndices(self.pose, self.tick, p.max_depth, p.time_delta)
                self.model.clean(self.pose, self.tick, p.confidence, p.time_delta)
        self._predict()
        self.pred.fillIn(filtered, rgb)
        self.poses.append(self.pose.copy())
        return self.pose
