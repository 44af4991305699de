"""Python mirror of the reference's odometry object over the C ABI.

``RGBDOdometry`` has the method names, argument order, defaults and error
behaviour of ``RGBDOdometryef`` (``src/odom/RGBDOdometryef.h:28-131``); the
GL textures of the reference become numpy arrays (host memory) or torch
tensors (device memory).  Every call goes through ``libslam_odom.so``; there
is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native

ODOM_OK = 0
ODOM_ERR_ARG = -1
ODOM_ERR_HIP = -2
ODOM_ERR_STATE = -3
ODOM_ERR_NODEV = -4

FMT_DEPTH_U16 = 1
FMT_RGB8 = 2
FMT_RGBA8 = 3
FMT_RGBA32F = 4

MEM_HOST = 0
MEM_DEVICE = 1

# Loop-closure acceptance thresholds of the reference's configs
# (src/configs/ef_file.cfg:9-14): icpErrThresh, covThresh, icpCountThresh.
ICP_ERR_THRESH = 5e-5
COV_THRESH = 1e-5
ICP_COUNT_THRESH = 40000


class FrameView(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("pitch_bytes", ctypes.c_int), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("format", ctypes.c_int), ("mem", ctypes.c_int)]


class Params(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("dist_thresh", ctypes.c_float),
                ("angle_thresh", ctypes.c_float)]


class Stats(ctypes.Structure):
    _fields_ = [("last_icp_error", ctypes.c_float), ("last_icp_count", ctypes.c_float),
                ("last_rgb_error", ctypes.c_float), ("last_rgb_count", ctypes.c_float),
                ("last_so3_error", ctypes.c_float), ("last_so3_count", ctypes.c_float),
                ("so3_iterations", ctypes.c_int), ("gn_iterations", ctypes.c_int),
                ("degenerate", ctypes.c_int), ("reverted", ctypes.c_int), ("seed_used", ctypes.c_int),
                ("seed_calls", ctypes.c_int), ("seed_kept", ctypes.c_int), ("alt_iterations", ctypes.c_int)]


START_PREVIOUS, START_VELOCITY = 0, 1
JUDGE_NONE, JUDGE_COARSE, JUDGE_FULL = 0, 1, 2


class Policy(ctypes.Structure):
    """odom_policy: schedule, convergence test, start and judge of a call."""
    _fields_ = [("iters", ctypes.c_int * 3), ("conv_t", ctypes.c_float), ("conv_r", ctypes.c_float),
                ("start", ctypes.c_int), ("so3", ctypes.c_int), ("judge", ctypes.c_int),
                ("judge_tol", ctypes.c_float)]

    @classmethod
    def make(cls, iters=(10, 5, 4), conv_t=0.0, conv_r=0.0, start=START_PREVIOUS, so3=False, judge=JUDGE_FULL,
             judge_tol=0.01) -> "Policy":
        p = cls()
        p.iters[:] = [int(v) for v in iters]
        p.conv_t, p.conv_r = float(conv_t), float(conv_r)
        p.start, p.so3, p.judge, p.judge_tol = int(start), int(so3), int(judge), float(judge_tol)
        return p


# Every symbol include/slam_odom.h declares.
EXPORTS = [
    "odom_default_params", "odom_create", "odom_destroy", "odom_bilateral_depth", "odom_init_icp_depth",
    "odom_init_icp_raw_depth", "odom_init_icp_maps", "odom_init_icp_model", "odom_init_rgb",
    "odom_init_rgb_model", "odom_init_first_rgb", "odom_track", "odom_track_async", "odom_set_pose",
    "odom_get_pose", "odom_synchronize", "odom_get_covariance", "odom_get_last_system", "odom_get_stats",
    "odom_icp_step", "odom_compute_derivatives", "odom_rgb_residual", "odom_rgb_step", "odom_so3_step",
    "odom_read_buffer", "odom_set_step_shape", "odom_get_step_shape", "odom_autotune", "odom_device_info",
    "odom_init_icp_model_current", "odom_profile_track", "odom_set_pose_log",
    "odom_batch_create", "odom_batch_destroy", "odom_track_batch", "odom_batch_synchronize",
    "odom_batch_set_blocks", "odom_batch_get_blocks", "odom_batch_set_step_shape", "odom_batch_set_graph",
    "odom_batch_profile_track", "odom_batch_set_split_solve", "odom_batch_autotune", "odom_batch_get_step_shape",
    "odom_get_stream", "odom_init_warmstart_input", "odom_track_seeded", "odom_set_graph", "odom_profile_steps",
    "odom_loop_check", "odom_icp_step_noalloc", "odom_default_policy", "odom_track_policy",
    "odom_batch_get_stream", "odom_batch_warmstart_input", "odom_track_batch_policy", "odom_stream_copy",
    "odom_set_so3_loop",
]

_V = ctypes.c_void_p
_P = ctypes.POINTER


def lib() -> ctypes.CDLL:
    L = _native.load("odom")
    if getattr(L, "_typed", False):
        return L
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    L.odom_default_params.argtypes = [_P(Params), ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 4
    L.odom_create.argtypes = [_P(Params), _P(_V)]
    L.odom_destroy.argtypes = [_V]
    L.odom_bilateral_depth.argtypes = [_V, _P(FrameView), ctypes.c_float, _V, ctypes.c_int]
    L.odom_init_icp_depth.argtypes = [_V, _P(FrameView), ctypes.c_float]
    L.odom_init_icp_raw_depth.argtypes = [_V, _P(FrameView), ctypes.c_float, ctypes.c_float]
    L.odom_init_icp_maps.argtypes = [_V, _P(FrameView), _P(FrameView), ctypes.c_float]
    L.odom_init_icp_model.argtypes = [_V, _P(FrameView), _P(FrameView), ctypes.c_float, _V]
    for n in ("odom_init_rgb", "odom_init_rgb_model", "odom_init_first_rgb"):
        getattr(L, n).argtypes = [_V, _P(FrameView)]
    L.odom_track.argtypes = [_V, _V, _V, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.odom_track_async.argtypes = [_V, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.odom_set_pose.argtypes = [_V, _V, _V]
    L.odom_set_so3_loop.argtypes = [_V, ctypes.c_int]
    L.odom_get_pose.argtypes = [_V, _V, _V]
    L.odom_synchronize.argtypes = [_V]
    L.odom_get_covariance.argtypes = [_V, _V]
    L.odom_loop_check.argtypes = [_V, ctypes.c_double, ctypes.c_float, ctypes.c_float, _P(ctypes.c_int)]
    L.odom_get_last_system.argtypes = [_V, _V, _V]
    L.odom_get_stats.argtypes = [_V, _P(Stats)]
    L.odom_icp_step.argtypes = [_V, ctypes.c_int, _V, _V, _V, _V, _V, _V]
    L.odom_compute_derivatives.argtypes = [_V]
    L.odom_rgb_residual.argtypes = [_V, ctypes.c_int, _V, _V, _V, _V, _V]
    L.odom_rgb_step.argtypes = [_V, ctypes.c_int, ctypes.c_float, _V, _V, _V]
    L.odom_so3_step.argtypes = [_V, _V, _V, _V, _V, _V]
    L.odom_read_buffer.argtypes = [_V, ctypes.c_char_p, ctypes.c_int, _V, ctypes.c_size_t]
    L.odom_set_step_shape.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.odom_get_step_shape.argtypes = [_V, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int)]
    L.odom_autotune.argtypes = [_V, ctypes.c_int, ctypes.c_float]
    L.odom_device_info.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.odom_init_icp_model_current.argtypes = [_V, _P(FrameView), _P(FrameView), ctypes.c_float]
    L.odom_profile_track.argtypes = [_V, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     _V, _V, _V, _V, _P(ctypes.c_float)]
    L.odom_set_pose_log.argtypes = [_V, _V, ctypes.c_int]
    L.odom_batch_create.argtypes = [_V, ctypes.c_int, _P(_V)]
    L.odom_batch_destroy.argtypes = [_V]
    L.odom_track_batch.argtypes = [_V, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.odom_batch_synchronize.argtypes = [_V]
    L.odom_batch_set_blocks.argtypes = [_V, ctypes.c_int]
    L.odom_batch_get_blocks.argtypes = [_V, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int)]
    L.odom_batch_set_step_shape.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.odom_batch_set_graph.argtypes = [_V, ctypes.c_int]
    L.odom_batch_set_split_solve.argtypes = [_V, ctypes.c_int]
    L.odom_get_stream.argtypes = [_V, _P(_V)]
    L.odom_set_graph.argtypes = [_V, ctypes.c_int]
    L.odom_profile_steps.argtypes = [_V, ctypes.c_int, ctypes.c_float, ctypes.c_int, _V, _V]
    L.odom_init_warmstart_input.argtypes = [_V, _V]
    L.odom_track_seeded.argtypes = [_V, _V, ctypes.c_float, ctypes.c_float, _V, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int]
    L.odom_batch_autotune.argtypes = [_V, ctypes.c_int, ctypes.c_float]
    L.odom_batch_get_step_shape.argtypes = [_V, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int)]
    L.odom_batch_profile_track.argtypes = [_V, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _V, _V, _V, _V, _P(ctypes.c_float)]
    L.odom_icp_step_noalloc.argtypes = [_V, _V, _V, _V, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 4 + \
        [_V, _V, _V, _V, ctypes.c_float, ctypes.c_float, _V, _V]
    L.odom_default_policy.argtypes = [_P(Policy)]
    L.odom_stream_copy.argtypes = [_V, _V, ctypes.c_size_t, ctypes.c_int, _P(ctypes.c_float), _V]
    L.odom_track_policy.argtypes = [_V, _V, ctypes.c_float, ctypes.c_float, _P(Policy), ctypes.c_int, ctypes.c_float]
    L.odom_batch_get_stream.argtypes = [_V, _P(_V)]
    L.odom_batch_warmstart_input.argtypes = [_V, _V]
    L.odom_track_batch_policy.argtypes = [_V, _V, ctypes.c_float, ctypes.c_float, _P(Policy), ctypes.c_int,
                                          ctypes.c_float]
    L._typed = True
    return L


class OdomError(RuntimeError):
    def __init__(self, fn: str, code: int):
        names = {ODOM_ERR_ARG: "ODOM_ERR_ARG", ODOM_ERR_HIP: "ODOM_ERR_HIP", ODOM_ERR_STATE: "ODOM_ERR_STATE",
                 ODOM_ERR_NODEV: "ODOM_ERR_NODEV"}
        super().__init__(f"{fn} returned {names.get(code, code)}")
        self.code = code


def _check(fn: str, rc: int) -> None:
    if rc != ODOM_OK:
        raise OdomError(fn, rc)


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


_FMT_BPP = {FMT_DEPTH_U16: 2, FMT_RGB8: 3, FMT_RGBA8: 4, FMT_RGBA32F: 16}


class DeviceImage:
    """A width x height image at a device address owned by a library handle
    (for instance a splat prediction's output): passed as an ODOM_MEM_DEVICE
    view.  ``owner`` keeps that handle alive."""

    def __init__(self, ptr: int, width: int, height: int, owner=None):
        self.ptr, self.width, self.height, self.owner = int(ptr), int(width), int(height), owner
        self.shape = (self.height, self.width)


def frame_view(img, fmt: int) -> tuple[FrameView, object]:
    """A view of a numpy array (host), a torch CUDA tensor or a DeviceImage
    (device).  Returns the view and the object that keeps its memory alive."""
    if isinstance(img, DeviceImage):
        return FrameView(img.ptr, img.width * _FMT_BPP[fmt], img.width, img.height, fmt, MEM_DEVICE), img
    try:
        import torch
        is_t = isinstance(img, torch.Tensor)
    except ImportError:  # pragma: no cover
        is_t = False
    if is_t:
        if not img.is_cuda:
            img = img.numpy()
        else:
            t = img.contiguous()
            H, W = t.shape[0], t.shape[1]
            v = FrameView(t.data_ptr(), W * _FMT_BPP[fmt], W, H, fmt, MEM_DEVICE)
            return v, t
    a = np.ascontiguousarray(img)
    H, W = a.shape[0], a.shape[1]
    v = FrameView(a.ctypes.data, W * _FMT_BPP[fmt], W, H, fmt, MEM_HOST)
    return v, a


def _rgb_fmt(img) -> int:
    if isinstance(img, DeviceImage):
        return FMT_RGBA8
    return FMT_RGB8 if img.shape[2] == 3 else FMT_RGBA8


class RGBDOdometry:
    """``RGBDOdometryef`` over ``libslam_odom.so``."""

    NUM_PYRS = 3

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 distThresh: float = 0.10, angleThresh: float | None = None):
        if angleThresh is None:
            angleThresh = math.sin(20.0 * 3.14159254 / 180.0)
        L = lib()
        p = Params()
        _check("odom_default_params", L.odom_default_params(ctypes.byref(p), width, height, cx, cy, fx, fy))
        p.dist_thresh = distThresh
        p.angle_thresh = angleThresh
        h = _V()
        _check("odom_create", L.odom_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.width, self.height = width, height

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().odom_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---------------------------------------------------------- init calls
    def bilateral(self, raw_depth, maxD: float) -> np.ndarray:
        v, keep = frame_view(raw_depth, FMT_DEPTH_U16)
        out = np.empty((self.height, self.width), np.uint16)
        _check("odom_bilateral_depth", lib().odom_bilateral_depth(self._h, ctypes.byref(v), maxD, _p(out), MEM_HOST))
        return out

    def initICP(self, filteredDepth, depthCutoff: float) -> None:
        v, keep = frame_view(filteredDepth, FMT_DEPTH_U16)
        _check("odom_init_icp_depth", lib().odom_init_icp_depth(self._h, ctypes.byref(v), depthCutoff))
        self._keep_icp = keep

    def initICPRaw(self, rawDepth, maxD: float, depthCutoff: float) -> None:
        v, keep = frame_view(rawDepth, FMT_DEPTH_U16)
        _check("odom_init_icp_raw_depth", lib().odom_init_icp_raw_depth(self._h, ctypes.byref(v), maxD, depthCutoff))
        self._keep_icp = keep

    def initICPMaps(self, predictedVertices, predictedNormals, depthCutoff: float) -> None:
        v, kv = frame_view(predictedVertices, FMT_RGBA32F)
        n, kn = frame_view(predictedNormals, FMT_RGBA32F)
        _check("odom_init_icp_maps", lib().odom_init_icp_maps(self._h, ctypes.byref(v), ctypes.byref(n), depthCutoff))
        self._keep_maps = (kv, kn)

    def initICPModel(self, predictedVertices, predictedNormals, depthCutoff: float, modelPose) -> None:
        v, kv = frame_view(predictedVertices, FMT_RGBA32F)
        n, kn = frame_view(predictedNormals, FMT_RGBA32F)
        pose = np.ascontiguousarray(modelPose, dtype=np.float32).reshape(16)
        _check("odom_init_icp_model",
               lib().odom_init_icp_model(self._h, ctypes.byref(v), ctypes.byref(n), depthCutoff, _p(pose)))
        self._keep_model = (kv, kn)

    def initICPModelCurrent(self, predictedVertices, predictedNormals, depthCutoff: float) -> None:
        """initICPModel with the device-resident current pose as modelPose."""
        v, kv = frame_view(predictedVertices, FMT_RGBA32F)
        n, kn = frame_view(predictedNormals, FMT_RGBA32F)
        _check("odom_init_icp_model_current",
               lib().odom_init_icp_model_current(self._h, ctypes.byref(v), ctypes.byref(n), depthCutoff))
        self._keep_model = (kv, kn)

    def initRGB(self, rgb) -> None:
        v, k = frame_view(rgb, _rgb_fmt(rgb))
        _check("odom_init_rgb", lib().odom_init_rgb(self._h, ctypes.byref(v)))
        self._keep_rgb = k

    def initRGBModel(self, rgb) -> None:
        v, k = frame_view(rgb, _rgb_fmt(rgb))
        _check("odom_init_rgb_model", lib().odom_init_rgb_model(self._h, ctypes.byref(v)))
        self._keep_rgb_model = k

    def initFirstRGB(self, rgb) -> None:
        v, k = frame_view(rgb, _rgb_fmt(rgb))
        _check("odom_init_first_rgb", lib().odom_init_first_rgb(self._h, ctypes.byref(v)))
        self._keep_first = k

    # ------------------------------------------------------------ tracking
    def getIncrementalTransformation(self, trans, rot, rgbOnly: bool = False, icpWeight: float = 10.0,
                                     pyramid: bool = True, fastOdom: bool = False, so3: bool = True):
        t = np.ascontiguousarray(trans, dtype=np.float32).reshape(3).copy()
        R = np.ascontiguousarray(rot, dtype=np.float32).reshape(9).copy()
        _check("odom_track", lib().odom_track(self._h, _p(t), _p(R), int(rgbOnly), float(icpWeight), int(pyramid),
                                              int(fastOdom), int(so3)))
        return t, R.reshape(3, 3)

    def set_graph(self, enable: bool) -> None:
        """Replay tracking calls from captured hipGraphs."""
        _check("odom_set_graph", lib().odom_set_graph(self._h, int(enable)))

    @property
    def stream(self) -> int:
        """The handle's HIP stream (as an integer pointer)."""
        st = _V()
        _check("odom_get_stream", lib().odom_get_stream(self._h, ctypes.byref(st)))
        return st.value or 0

    def initWarmStartInput(self, x) -> None:
        """Enqueue the warm-start network input (torch bf16 [1][60][80][4] on
        the GPU) from this handle's model and current level-2 buffers."""
        _check("odom_init_warmstart_input", lib().odom_init_warmstart_input(self._h, x.data_ptr()))

    def track_seeded(self, pred, iters=(5, 2, 2), rgbOnly=False, icpWeight=10.0, select=True,
                     t_scale=0.05, r_scale=0.05) -> None:
        """Asynchronous tracking call starting from a predicted motion (device
        tensor of 6 floats), with an explicit schedule and no SO3 pass."""
        it = np.ascontiguousarray(iters, dtype=np.int32)
        _check("odom_track_seeded", lib().odom_track_seeded(self._h, pred.data_ptr(), float(t_scale), float(r_scale),
                                                            it.ctypes.data, int(rgbOnly), float(icpWeight),
                                                            int(select)))

    def track_policy(self, pred=None, policy: Policy | None = None, rgbOnly=False, icpWeight=10.0, t_scale=0.05,
                     r_scale=0.05) -> None:
        """Asynchronous tracking call under a policy (odom_track_policy):
        pred is a device tensor of 6 floats or None."""
        pol = policy if policy is not None else Policy.make()
        ptr = None if pred is None else pred.data_ptr()
        _check("odom_track_policy", lib().odom_track_policy(self._h, ptr, float(t_scale), float(r_scale),
                                                            ctypes.byref(pol), int(rgbOnly), float(icpWeight)))

    def track_async(self, rgbOnly=False, icpWeight=10.0, pyramid=True, fastOdom=False, so3=True) -> None:
        _check("odom_track_async", lib().odom_track_async(self._h, int(rgbOnly), float(icpWeight), int(pyramid),
                                                          int(fastOdom), int(so3)))

    def set_pose(self, trans, rot) -> None:
        t = np.ascontiguousarray(trans, dtype=np.float32).reshape(3)
        R = np.ascontiguousarray(rot, dtype=np.float32).reshape(9)
        _check("odom_set_pose", lib().odom_set_pose(self._h, _p(t), _p(R)))

    def get_pose(self):
        t = np.zeros(3, np.float32)
        R = np.zeros(9, np.float32)
        _check("odom_get_pose", lib().odom_get_pose(self._h, _p(t), _p(R)))
        return t, R.reshape(3, 3)

    def set_pose_log(self, dev_ptr: int, capacity: int) -> None:
        _check("odom_set_pose_log", lib().odom_set_pose_log(self._h, dev_ptr, capacity))

    def synchronize(self) -> None:
        _check("odom_synchronize", lib().odom_synchronize(self._h))

    def getCovariance(self) -> np.ndarray:
        c = np.zeros(36, np.float64)
        _check("odom_get_covariance", lib().odom_get_covariance(self._h, _p(c)))
        return c.reshape(6, 6)

    def loopCheck(self, covThresh: float, icpCountThresh: float, icpErrThresh: float) -> bool:
        """The frame loop's loop-closure acceptance test on the last call
        (elastic_fusion_file.cpp:474-497)."""
        ok = ctypes.c_int(0)
        _check("odom_loop_check", lib().odom_loop_check(self._h, float(covThresh), float(icpCountThresh),
                                                        float(icpErrThresh), ctypes.byref(ok)))
        return bool(ok.value)

    def set_so3_loop(self, spin_limit: int = 1 << 20) -> None:
        """The SO3 loop as one launch with a bounded grid barrier (spin_limit
        polls per wait), or one launch per iteration (spin_limit < 0)."""
        _check("odom_set_so3_loop", lib().odom_set_so3_loop(self._h, int(spin_limit)))

    def modelToModel(self, inactiveVertices, inactiveNormals, inactiveImage, activeVertices, activeNormals,
                     activeImage, pose, depthCutoff: float, pyramid: bool = True, fastOdom: bool = False,
                     covThresh: float = COV_THRESH, icpCountThresh: float = ICP_COUNT_THRESH,
                     icpErrThresh: float = ICP_ERR_THRESH):
        """The frame loop's local loop-closure check on this handle (its
        ``modelToModel`` odometry), elastic_fusion_file.cpp:449-498: the
        active model prediction is registered to the inactive one from
        ``pose`` (4x4) with ICP and RGB and SO3 off, and accepted when every
        diagonal entry of the covariance (inverse of the last 6x6 system) is
        at most ``covThresh``, more than ``icpCountThresh`` pixels were ICP
        inliers and the ICP error is below ``icpErrThresh`` (odom_loop_check).

        The vertex/normal maps are (h, w, 4) float32 in the camera frame of
        ``pose``, the images (h, w, 3 or 4) uint8, host arrays or device
        tensors.  Returns (accepted, estPose) with estPose the registered 4x4
        float32 estimate."""
        pose = np.asarray(pose, dtype=np.float32)
        # initICP* before initRGB*: the RGB inits read the vertex maps
        self.initICPModel(inactiveVertices, inactiveNormals, depthCutoff, pose)
        self.initRGBModel(inactiveImage)
        self.initICPMaps(activeVertices, activeNormals, depthCutoff)
        self.initRGB(activeImage)
        t, R = self.getIncrementalTransformation(pose[:3, 3], pose[:3, :3], False, 10.0, pyramid, fastOdom, False)
        est = np.eye(4, dtype=np.float32)
        est[:3, :3] = np.asarray(R, np.float32).reshape(3, 3)
        est[:3, 3] = np.asarray(t, np.float32).reshape(3)
        return self.loopCheck(covThresh, icpCountThresh, icpErrThresh), est

    def stats(self) -> dict:
        s = Stats()
        _check("odom_get_stats", lib().odom_get_stats(self._h, ctypes.byref(s)))
        return {"lastICPError": s.last_icp_error, "lastICPCount": s.last_icp_count,
                "lastRGBError": s.last_rgb_error, "lastRGBCount": s.last_rgb_count,
                "lastSO3Error": s.last_so3_error, "lastSO3Count": s.last_so3_count,
                "so3Iterations": s.so3_iterations, "gnIterations": s.gn_iterations,
                "degenerate": s.degenerate, "reverted": s.reverted, "seedUsed": s.seed_used,
                "seedCalls": s.seed_calls, "seedKept": s.seed_kept, "altIterations": s.alt_iterations}

    @property
    def lastA(self) -> np.ndarray:
        A = np.zeros(36, np.float64)
        b = np.zeros(6, np.float64)
        _check("odom_get_last_system", lib().odom_get_last_system(self._h, _p(A), _p(b)))
        return A.reshape(6, 6)

    @property
    def lastb(self) -> np.ndarray:
        A = np.zeros(36, np.float64)
        b = np.zeros(6, np.float64)
        _check("odom_get_last_system", lib().odom_get_last_system(self._h, _p(A), _p(b)))
        return b

    # ------------------------------------------------------ kernel level
    def icp_step(self, level, Rcurr, tcurr, Rprev_inv, tprev, mask=True):
        a = [np.ascontiguousarray(x, dtype=np.float32).reshape(-1) for x in (Rcurr, tcurr, Rprev_inv, tprev)]
        out = np.zeros(29, np.float64)
        m = np.zeros((self.height >> level, self.width >> level), np.uint8) if mask else None
        _check("odom_icp_step", lib().odom_icp_step(self._h, level, *[_p(x) for x in a], _p(out),
                                                    _p(m) if mask else None))
        return out, m

    def compute_derivatives(self) -> None:
        _check("odom_compute_derivatives", lib().odom_compute_derivatives(self._h))

    def rgb_residual(self, level, krkinv, kt, mask=True):
        K = np.ascontiguousarray(krkinv, dtype=np.float32).reshape(9)
        t = np.ascontiguousarray(kt, dtype=np.float32).reshape(3)
        cnt = np.zeros(1, np.int32)
        sig = np.zeros(1, np.int32)
        m = np.zeros((self.height >> level, self.width >> level), np.uint8) if mask else None
        _check("odom_rgb_residual", lib().odom_rgb_residual(self._h, level, _p(K), _p(t), _p(cnt), _p(sig),
                                                            _p(m) if mask else None))
        return int(cnt[0]), int(sig[0]), m

    def rgb_step(self, level, sigma, krkinv, kt):
        K = np.ascontiguousarray(krkinv, dtype=np.float32).reshape(9)
        t = np.ascontiguousarray(kt, dtype=np.float32).reshape(3)
        out = np.zeros(29, np.float64)
        _check("odom_rgb_step", lib().odom_rgb_step(self._h, level, float(sigma), _p(K), _p(t), _p(out)))
        return out

    def so3_step(self, H, kinv, krlr, mask=True):
        a = [np.ascontiguousarray(x, dtype=np.float32).reshape(9) for x in (H, kinv, krlr)]
        out = np.zeros(11, np.float64)
        m = np.zeros((self.height >> 2, self.width >> 2), np.uint8) if mask else None
        _check("odom_so3_step", lib().odom_so3_step(self._h, *[_p(x) for x in a], _p(out), _p(m) if mask else None))
        return out, m

    _BUF = {"depth": (np.uint16, 1), "vmap_curr": (np.float32, 3), "nmap_curr": (np.float32, 3),
            "vmap_g_prev": (np.float32, 3), "nmap_g_prev": (np.float32, 3), "last_depth": (np.float32, 1),
            "next_depth": (np.float32, 1), "last_image": (np.uint8, 1), "next_image": (np.uint8, 1),
            "last_next_image": (np.uint8, 1), "dxy": (np.int16, 2), "zsrc": (np.float32, 1)}

    def read_buffer(self, name: str, level: int) -> np.ndarray:
        dt, planes = self._BUF[name]
        h, w = self.height >> level, self.width >> level
        if name == "dxy":
            out = np.zeros((h, w, 2), dt)
        elif planes == 3:
            out = np.zeros((3, h, w), dt)
        else:
            out = np.zeros((h, w), dt)
        _check("odom_read_buffer", lib().odom_read_buffer(self._h, name.encode(), level, _p(out), out.nbytes))
        return out

    def set_step_shape(self, kernel: int, level: int, threads: int, ppt: int) -> None:
        _check("odom_set_step_shape", lib().odom_set_step_shape(self._h, kernel, level, threads, ppt))

    def get_step_shape(self, kernel: int, level: int) -> tuple[int, int]:
        t, p = ctypes.c_int(), ctypes.c_int()
        _check("odom_get_step_shape", lib().odom_get_step_shape(self._h, kernel, level, ctypes.byref(t),
                                                                ctypes.byref(p)))
        return t.value, p.value

    def profile_track(self, rgbOnly=False, icpWeight=10.0, pyramid=True, fastOdom=False, so3=True) -> dict:
        """One tracking call with HIP events around every step launch."""
        ua, ub = np.zeros(3, np.float32), np.zeros(3, np.float32)
        na, nb = np.zeros(3, np.int32), np.zeros(3, np.int32)
        tot = ctypes.c_float()
        _check("odom_profile_track", lib().odom_profile_track(self._h, int(rgbOnly), float(icpWeight), int(pyramid),
                                                              int(fastOdom), int(so3), _p(ua), _p(ub), _p(na),
                                                              _p(nb), ctypes.byref(tot)))
        return {"us_a": ua.tolist(), "us_b": ub.tolist(), "n_a": na.tolist(), "n_b": nb.tolist(),
                "us_total": tot.value}

    def profile_steps(self, rgbOnly=False, icpWeight=10.0, reps=50) -> dict:
        """Mean microseconds per launch of the two step kernels per level, as
        they run inside a tracking call (needs a completed call)."""
        ua, ub = np.zeros(3, np.float32), np.zeros(3, np.float32)
        _check("odom_profile_steps", lib().odom_profile_steps(self._h, int(rgbOnly), float(icpWeight), int(reps),
                                                              _p(ua), _p(ub)))
        return {"us_a": ua.tolist(), "us_b": ub.tolist()}

    def autotune(self, rgbOnly=False, icpWeight=10.0) -> None:
        _check("odom_autotune", lib().odom_autotune(self._h, int(rgbOnly), float(icpWeight)))


class Batch:
    """Many RGBDOdometry handles (frame pairs of any size) tracked by the same
    launches (``odom_track_batch``).  Inputs go to each handle with its own
    init* calls; results are read per handle."""

    def __init__(self, odos):
        self.odos = list(odos)
        arr = (_V * len(self.odos))(*[o.handle for o in self.odos])
        b = _V()
        _check("odom_batch_create", lib().odom_batch_create(arr, len(self.odos), ctypes.byref(b)))
        self._b = b

    def close(self) -> None:
        if getattr(self, "_b", None):
            lib().odom_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def track(self, rgbOnly=False, icpWeight=10.0, pyramid=True, fastOdom=False, so3=True) -> None:
        _check("odom_track_batch", lib().odom_track_batch(self._b, int(rgbOnly), float(icpWeight), int(pyramid),
                                                          int(fastOdom), int(so3)))

    def synchronize(self) -> None:
        _check("odom_batch_synchronize", lib().odom_batch_synchronize(self._b))

    def track_policy(self, pred=None, policy: Policy | None = None, rgbOnly=False, icpWeight=10.0, t_scale=0.05,
                     r_scale=0.05) -> None:
        """odom_track_batch_policy: pred is a device tensor [n][6] or None."""
        pol = policy if policy is not None else Policy.make()
        ptr = None if pred is None else pred.data_ptr()
        _check("odom_track_batch_policy", lib().odom_track_batch_policy(self._b, ptr, float(t_scale), float(r_scale),
                                                                        ctypes.byref(pol), int(rgbOnly),
                                                                        float(icpWeight)))

    def warmstart_input(self, x) -> None:
        """The network input of every pair into x (bf16 [n][60][80][4])."""
        _check("odom_batch_warmstart_input", lib().odom_batch_warmstart_input(self._b, x.data_ptr()))

    @property
    def stream(self) -> int:
        """The batch's HIP stream (as an integer pointer)."""
        st = _V()
        _check("odom_batch_get_stream", lib().odom_batch_get_stream(self._b, ctypes.byref(st)))
        return st.value or 0

    def set_blocks(self, target_blocks: int) -> None:
        _check("odom_batch_set_blocks", lib().odom_batch_set_blocks(self._b, int(target_blocks)))

    def get_blocks(self, kernel: int, level: int) -> int:
        n = ctypes.c_int()
        _check("odom_batch_get_blocks", lib().odom_batch_get_blocks(self._b, kernel, level, ctypes.byref(n)))
        return n.value

    def set_step_shape(self, kernel: int, level: int, threads: int, ppt: int) -> None:
        _check("odom_batch_set_step_shape", lib().odom_batch_set_step_shape(self._b, kernel, level, threads, ppt))

    def autotune(self, rgbOnly=False, icpWeight=10.0) -> None:
        _check("odom_batch_autotune", lib().odom_batch_autotune(self._b, int(rgbOnly), float(icpWeight)))

    def get_step_shape(self, kernel: int, level: int) -> tuple[int, int]:
        t, p = ctypes.c_int(), ctypes.c_int()
        _check("odom_batch_get_step_shape", lib().odom_batch_get_step_shape(self._b, kernel, level, ctypes.byref(t),
                                                                            ctypes.byref(p)))
        return t.value, p.value

    def set_split_solve(self, enable: bool) -> None:
        _check("odom_batch_set_split_solve", lib().odom_batch_set_split_solve(self._b, int(enable)))

    def set_graph(self, enable: bool) -> None:
        _check("odom_batch_set_graph", lib().odom_batch_set_graph(self._b, int(enable)))

    def profile_track(self, rgbOnly=False, icpWeight=10.0, pyramid=True, fastOdom=False, so3=True) -> dict:
        ua, ub = np.zeros(3, np.float32), np.zeros(3, np.float32)
        na, nb = np.zeros(3, np.int32), np.zeros(3, np.int32)
        tot = ctypes.c_float()
        _check("odom_batch_profile_track",
               lib().odom_batch_profile_track(self._b, int(rgbOnly), float(icpWeight), int(pyramid), int(fastOdom),
                                              int(so3), _p(ua), _p(ub), _p(na), _p(nb), ctypes.byref(tot)))
        return {"us_a": ua.tolist(), "us_b": ub.tolist(), "n_a": na.tolist(), "n_b": nb.tolist(),
                "us_total": tot.value}


def device_info() -> tuple[str, str]:
    n = ctypes.create_string_buffer(256)
    a = ctypes.create_string_buffer(128)
    _check("odom_device_info", lib().odom_device_info(n, 256, a, 128))
    return n.value.decode(), a.value.decode()
