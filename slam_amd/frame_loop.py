"""The reference's frame loop without loop closure, on the device: bilateral
filter, frame-to-model tracking against the filled-in splat prediction, the
model update, and the next prediction (elastic_fusion_file.cpp:335-420,
551-602, with closeLoops off).

``FrameLoop`` drives the native loop of ``include/slam_loop.h``
(``efloop_process``, C++ over the odometry, splat and surfel entry points).
``BatchFrameLoop`` runs n sequences in lockstep (``efloop_batch``), their
tracking calls batched and, with the learned start, one network forward for
all of them; a FrameLoop with any other tracking policy is a batch of one.
Every stage of a sequence runs on one HIP stream (its odometry's); the host
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
from .odometry import (FMT_DEPTH_U16, FMT_RGB8, JUDGE_FULL, START_PREVIOUS, Policy, Stats, _check, _p, frame_view,
                       lib as odom_lib)

# Every symbol include/slam_loop.h declares.
EXPORTS = ["efloop_default_params", "efloop_create", "efloop_destroy", "efloop_process", "efloop_get_handles",
           "efloop_get_state", "efloop_batch_create", "efloop_batch_destroy", "efloop_batch_set_net",
           "efloop_batch_process", "efloop_batch_get_loop", "efloop_batch_get_counters"]


class _Params(ctypes.Structure):
    _fields_ = [("depth_cutoff", ctypes.c_float), ("max_depth", ctypes.c_float), ("confidence", ctypes.c_float),
                ("time_delta", ctypes.c_int), ("icp_weight", ctypes.c_float), ("so3", ctypes.c_int),
                ("max_surfels", ctypes.c_int), ("use_graph", ctypes.c_int), ("autotune", ctypes.c_int)]


class Counters(ctypes.Structure):
    _fields_ = [("frames_tracked", ctypes.c_int), ("gn_iterations", ctypes.c_longlong),
                ("alt_iterations", ctypes.c_longlong), ("predictions", ctypes.c_int),
                ("predictions_kept", ctypes.c_int)]


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
    from .odometry import Policy
    L.efloop_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 4 + \
        [P(_Params), P(Policy), P(V)]
    L.efloop_batch_destroy.argtypes = [V]
    L.efloop_batch_set_net.argtypes = [V, V, ctypes.c_size_t, ctypes.c_float, ctypes.c_float]
    L.efloop_batch_process.argtypes = [V, V, V, V]
    L.efloop_batch_get_loop.argtypes = [V, ctypes.c_int, P(V)]
    L.efloop_batch_get_counters.argtypes = [V, ctypes.c_int, P(Counters)]
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
                "gnIterations": s.gn_iterations, "altIterations": s.alt_iterations,
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
    """The reference configs' values (src/configs/ef_file.cfg), and the
    tracking policy of the batched loop."""
    depth_cutoff: float = 3.0        # depthCutoff: bilateral and metric depth window (m)
    max_depth: float = 20.0          # maxDepthProcessed
    confidence: float = 10.0
    time_delta: int = 200
    icp_weight: float = 10.0
    so3: bool = True
    max_surfels: int = 4_000_000
    use_graph: bool = True           # replay the tracking call from a hipGraph
    autotune: bool = False           # step-kernel launch shapes timed on the first tracked frame (the
                                     # library's table otherwise; timing makes runs differ in their last bits)
    learned: bool = False            # start each call from the warm-start network's prediction, judged
                                     # against the constant-velocity start (no SO3 pass then)
    iters: tuple = (10, 5, 4)        # the policy's schedule (levels 0, 1, 2)
    conv_t: float = 0.0              # the policy's convergence test (m, rad); 0 = the fixed schedule
    conv_r: float = 0.0
    start: int = START_PREVIOUS      # the start without a prediction
    judge: int = JUDGE_FULL
    judge_tol: float = 0.01

    def policy(self) -> Policy:
        so3 = bool(self.so3) and not self.learned and self.start == START_PREVIOUS
        return Policy.make(self.iters, self.conv_t, self.conv_r, self.start, so3, self.judge, self.judge_tol)

    def is_reference(self) -> bool:
        """The reference's getIncrementalTransformation: previous-pose start,
        fixed 10/5/4 schedule, no prediction."""
        return (not self.learned and self.start == START_PREVIOUS and tuple(self.iters) == (10, 5, 4)
                and self.conv_t == 0.0)


def _native_params(p: LoopParams) -> _Params:
    P = _Params()
    _check("efloop_default_params", lib().efloop_default_params(ctypes.byref(P)))
    P.depth_cutoff, P.max_depth, P.confidence = p.depth_cutoff, p.max_depth, p.confidence
    P.time_delta, P.icp_weight, P.so3 = p.time_delta, p.icp_weight, int(p.so3)
    P.max_surfels = p.max_surfels
    P.use_graph = int(p.use_graph)
    P.autotune = int(p.autotune)
    return P


class FrameLoop:
    """Processes frames (uint16 depth in mm, RGB8), host or device.  The
    reference's loop runs natively per frame (efloop_process); any other
    tracking policy (the learned start, the constant-velocity start, the
    convergence test) runs as a batched loop of one sequence."""

    def __new__(cls, width, height, cx, cy, fx, fy, params: LoopParams | None = None):
        if params is not None and not params.is_reference():
            return super().__new__(_BatchOfOne)
        return super().__new__(cls)

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 params: LoopParams | None = None):
        self.p = params or LoopParams()
        P = _native_params(self.p)
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


class BatchFrameLoop:
    """``n`` sequences of one size in lockstep (include/slam_loop.h
    efloop_batch): per frame one batched tracking call for all of them under
    the params' policy, with the learned start's network run once for the
    whole batch.  Each sequence's results equal those of the same sequence
    run alone."""

    def __init__(self, n: int, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 params: LoopParams | None = None, weights: str | None = None):
        self.p = params or LoopParams()
        self.n = n
        P = _native_params(self.p)
        pol = self.p.policy()
        h = ctypes.c_void_p()
        _check("efloop_batch_create", lib().efloop_batch_create(n, width, height, cx, cy, fx, fy, ctypes.byref(P),
                                                                ctypes.byref(pol), ctypes.byref(h)))
        self._h = h
        if self.p.learned:
            from . import warmstart as ws
            with np.load(weights or ws.WEIGHTS, allow_pickle=False) as z:
                params = np.ascontiguousarray(z["params"], dtype=np.float32)
                if float(z["t_scale"]) != ws.T_SCALE or float(z["r_scale"]) != ws.R_SCALE:
                    raise ValueError("weights were trained with other output scales")
            _check("efloop_batch_set_net", lib().efloop_batch_set_net(h, params.ctypes.data, params.size, ws.T_SCALE,
                                                                      ws.R_SCALE))
        self.odos, self.models = [], []
        for i in range(n):
            loop = ctypes.c_void_p()
            _check("efloop_batch_get_loop", lib().efloop_batch_get_loop(h, i, ctypes.byref(loop)))
            o, m = ctypes.c_void_p(), ctypes.c_void_p()
            _check("efloop_get_handles", lib().efloop_get_handles(loop, ctypes.byref(o), None, ctypes.byref(m)))
            self.odos.append(_OdoView(o))
            self.models.append(_ModelView(m))
        self.poses: list[list[np.ndarray]] = [[] for _ in range(n)]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().efloop_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass

    def process(self, rgbs, depths) -> list[np.ndarray]:
        """One frame of every sequence; returns the n tracked poses (4x4)."""
        FV = odom_frame_view_type()
        rv, dv, keep = (FV * self.n)(), (FV * self.n)(), []
        for i in range(self.n):
            r, kr = frame_view(rgbs[i], FMT_RGB8)
            d, kd = frame_view(depths[i], FMT_DEPTH_U16)
            rv[i], dv[i] = r, d
            keep += [kr, kd]
        out = np.zeros((self.n, 16), np.float32)
        _check("efloop_batch_process", lib().efloop_batch_process(self._h, rv, dv, _p(out)))
        poses = [out[i].reshape(4, 4).copy() for i in range(self.n)]
        for i, T in enumerate(poses):
            self.poses[i].append(T)
        return poses

    def counters(self, i: int) -> dict:
        c = Counters()
        _check("efloop_batch_get_counters", lib().efloop_batch_get_counters(self._h, i, ctypes.byref(c)))
        return {"frames_tracked": c.frames_tracked, "gn_iterations": c.gn_iterations,
                "alt_iterations": c.alt_iterations, "predictions": c.predictions,
                "predictions_kept": c.predictions_kept}


class _BatchOfOne(FrameLoop):
    """FrameLoop under a non-reference policy: a batched loop of one."""

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 params: LoopParams | None = None):
        self.p = params or LoopParams()
        self._b = BatchFrameLoop(1, width, height, cx, cy, fx, fy, self.p)
        self._h = None
        self.odo = self._b.odos[0]
        self.model = self._b.models[0]
        self.poses = self._b.poses[0]

    def close(self) -> None:
        if getattr(self, "_b", None) is not None:
            self._b.close()
            self._b = None

    def process(self, rgb, depth) -> np.ndarray:
        return self._b.process([rgb], [depth])[0]

    def counters(self) -> dict:
        return self._b.counters(0)
