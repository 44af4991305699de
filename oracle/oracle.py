"""ctypes bindings of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY.  This is the checker of the HIP product path; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import it.  See ``odom_oracle.h`` for what each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_U16 = ctypes.POINTER(ctypes.c_uint16)
_S16 = ctypes.POINTER(ctypes.c_int16)
_I32 = ctypes.POINTER(ctypes.c_int32)
_V = ctypes.c_void_p


class Intr(ctypes.Structure):
    _fields_ = [("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float)]


DATATERM = np.dtype([("zero_x", "<i2"), ("zero_y", "<i2"), ("one_x", "<i2"), ("one_y", "<i2"),
                     ("diff", "<f4"), ("valid", "<i4")])

class SplatParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("depth_cutoff", ctypes.c_float),
                ("conf_threshold", ctypes.c_float), ("time", ctypes.c_int), ("max_time", ctypes.c_int),
                ("time_delta", ctypes.c_int)]


_P_SPLAT = ctypes.POINTER(SplatParams)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_level_intr.restype = Intr
        L.orc_level_intr.argtypes = [Intr, ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_hardware_concurrency.restype = ctypes.c_int
        L.orc_bexpf.restype = ctypes.c_float
        L.orc_bexpf.argtypes = [ctypes.c_float]
        L.orc_create.restype = _V
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 6
        L.orc_destroy.argtypes = [_V]
        L.orc_set_precision.argtypes = [_V, ctypes.c_int]
        L.orc_init_icp_depth.argtypes = [_V, _V, ctypes.c_float]
        L.orc_init_icp_maps.argtypes = [_V, _V, _V, ctypes.c_float]
        L.orc_init_icp_model.argtypes = [_V, _V, _V, ctypes.c_float, _V]
        L.orc_init_rgb.argtypes = [_V, _V, ctypes.c_int]
        L.orc_init_rgb_model.argtypes = [_V, _V, ctypes.c_int]
        L.orc_init_first_rgb.argtypes = [_V, _V, ctypes.c_int]
        L.orc_track.argtypes = [_V, _V, _V, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_get_stats.argtypes = [_V, _V]
        L.orc_get_last_system.argtypes = [_V, _V, _V]
        L.orc_get_covariance.argtypes = [_V, _V]
        L.orc_get_covariance.restype = ctypes.c_int
        L.orc_get_so3_rotation.argtypes = [_V, _V]
        L.orc_track_ex.argtypes = [_V, _V, _V, ctypes.c_int, ctypes.c_float, _V, ctypes.c_int, _V]
        L.orc_track_run.argtypes = [_V, _V, _V, ctypes.c_int, ctypes.c_float, _V, ctypes.c_int, _V, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_int, ctypes.c_int, _V, _V]
        for n, rt in (("orc_buf_f32", _F), ("orc_buf_u8", _U8), ("orc_buf_s16", _S16), ("orc_buf_u16", _U16)):
            getattr(L, n).restype = rt
            getattr(L, n).argtypes = [_V, ctypes.c_char_p, ctypes.c_int]
        L.orc_ldlt_solve.restype = ctypes.c_int
        L.orc_ldlt_solve.argtypes = [ctypes.c_int, _V, _V, _V]
        L.orc_rodrigues.argtypes = [_V, _V]
        L.orc_icp_step.argtypes = [_V] * 4 + [ctypes.c_int, ctypes.c_int, Intr] + [_V] * 4 + \
            [ctypes.c_float, ctypes.c_float, _V, _V]
        L.orc_icp_step_f64.argtypes = [_V] * 4 + [ctypes.c_int, ctypes.c_int, Intr] + [_V] * 4 + \
            [ctypes.c_double, ctypes.c_double, _V, _V]
        L.orc_rgb_residual.argtypes = [ctypes.c_float] + [_V] * 6 + [ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                                                     _V, _V, _V, _V, _V]
        L.orc_rgb_step.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_float, _V, Intr, _V, _V,
                                   ctypes.c_float, _V]
        L.orc_so3_step.argtypes = [_V, _V, ctypes.c_int, ctypes.c_int, _V, _V, _V, _V, _V]
        L.orc_unpack_se3.argtypes = [_V, _V, _V]
        L.orc_bilateral_u16.argtypes = [_V, _V, ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.orc_pyrdown_u16.argtypes = [_V, ctypes.c_int, ctypes.c_int, _V]
        L.orc_vmap.argtypes = [_V, ctypes.c_int, ctypes.c_int, Intr, ctypes.c_float, _V]
        L.orc_nmap.argtypes = [_V, ctypes.c_int, ctypes.c_int, _V]
        L.orc_copy_maps.argtypes = [_V, _V, ctypes.c_int, ctypes.c_int, _V, _V]
        L.orc_resize_map.argtypes = [_V, ctypes.c_int, ctypes.c_int, _V, ctypes.c_int]
        L.orc_transform_maps.argtypes = [_V, _V, ctypes.c_int, ctypes.c_int, _V, _V]
        L.orc_vertices_to_depth.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_float, _V]
        L.orc_pyrdown_f32.argtypes = [_V, ctypes.c_int, ctypes.c_int, _V]
        L.orc_intensity.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _V]
        L.orc_pyrdown_u8.argtypes = [_V, ctypes.c_int, ctypes.c_int, _V]
        L.orc_derivative.argtypes = [_V, ctypes.c_int, ctypes.c_int, _V, _V]
        L.orc_point_cloud.argtypes = [_V, ctypes.c_int, ctypes.c_int, Intr, _V]
        L.orc_splat_predict.argtypes = [_P_SPLAT, _V, _V, ctypes.c_int, _V, _V, _V, _V]
        _I, _Fl = ctypes.c_int, ctypes.c_float
        L.orc_model_create.restype = _V
        L.orc_model_create.argtypes = [_I, _I, _Fl, _Fl, _Fl, _Fl, _I]
        L.orc_model_destroy.argtypes = [_V]
        L.orc_model_count.argtypes = [_V]
        L.orc_model_fresh_count.argtypes = [_V]
        L.orc_model_data.restype = _V
        L.orc_model_data.argtypes = [_V]
        L.orc_model_fresh.restype = _V
        L.orc_model_fresh.argtypes = [_V]
        L.orc_model_set.argtypes = [_V, _V, _I]
        L.orc_model_index_map.argtypes = [_V, _V, _V, _V, _V]
        L.orc_depth_metric.argtypes = [_V, _I, _Fl, _V]
        L.orc_model_initialise.argtypes = [_V, _V, _V, _V, _Fl, _I, _Fl]
        L.orc_model_predict_indices.argtypes = [_V, _V, _I, _Fl, _I]
        L.orc_model_fuse.argtypes = [_V, _V, _I, _V, _V, _V, _Fl, _Fl, _Fl]
        L.orc_model_clean.argtypes = [_V, _V, _I, _Fl, _I]
        L.orc_splat_fill_in.argtypes = [_P_SPLAT, _V, _V, _V, _V, _V, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _V, _V, _V]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def intr(fx, fy, cx, cy) -> Intr:
    return Intr(fx, fy, cx, cy)


def level_intr(k: Intr, level: int) -> Intr:
    return lib().orc_level_intr(k, level)


# ------------------------------------------------------------------ image ops

def bilateral(depth, maxD):
    d = _c(depth, np.uint16)
    out = np.empty_like(d)
    lib().orc_bilateral_u16(_p(d), _p(out), d.shape[1], d.shape[0], maxD)
    return out


def pyrdown_u16(src):
    s = _c(src, np.uint16)
    out = np.empty((s.shape[0] // 2, s.shape[1] // 2), np.uint16)
    lib().orc_pyrdown_u16(_p(s), s.shape[1], s.shape[0], _p(out))
    return out


def vmap(depth, k: Intr, cutoff):
    d = _c(depth, np.uint16)
    out = np.empty((3,) + d.shape, np.float32)
    lib().orc_vmap(_p(d), d.shape[1], d.shape[0], k, cutoff, _p(out))
    return out


def nmap(vm):
    v = _c(vm, np.float32)
    out = np.empty_like(v)
    lib().orc_nmap(_p(v), v.shape[2], v.shape[1], _p(out))
    return out


def copy_maps(v4, n4):
    v, n = _c(v4, np.float32), _c(n4, np.float32)
    H, W = v.shape[:2]
    vo = np.empty((3, H, W), np.float32)
    no = np.empty((3, H, W), np.float32)
    lib().orc_copy_maps(_p(v), _p(n), W, H, _p(vo), _p(no))
    return vo, no


def resize_map(m, normalize):
    s = _c(m, np.float32)
    out = np.empty((3, s.shape[1] // 2, s.shape[2] // 2), np.float32)
    lib().orc_resize_map(_p(s), s.shape[2], s.shape[1], _p(out), int(normalize))
    return out


def transform_maps(vm, nm, R, t):
    v, n = _c(vm, np.float32).copy(), _c(nm, np.float32).copy()
    Rf, tf = _c(R, np.float32), _c(t, np.float32)
    lib().orc_transform_maps(_p(v), _p(n), v.shape[2], v.shape[1], _p(Rf), _p(tf))
    return v, n


def vertices_to_depth(v4, cutoff=6.0):
    v = _c(v4, np.float32)
    out = np.empty(v.shape[:2], np.float32)
    lib().orc_vertices_to_depth(_p(v), v.shape[1], v.shape[0], cutoff, _p(out))
    return out


def pyrdown_f32(src):
    s = _c(src, np.float32)
    out = np.empty((s.shape[0] // 2, s.shape[1] // 2), np.float32)
    lib().orc_pyrdown_f32(_p(s), s.shape[1], s.shape[0], _p(out))
    return out


def intensity(rgb):
    r = _c(rgb, np.uint8)
    out = np.empty(r.shape[:2], np.uint8)
    lib().orc_intensity(_p(r), r.shape[1], r.shape[0], r.shape[2], _p(out))
    return out


def pyrdown_u8(src):
    s = _c(src, np.uint8)
    out = np.empty((s.shape[0] // 2, s.shape[1] // 2), np.uint8)
    lib().orc_pyrdown_u8(_p(s), s.shape[1], s.shape[0], _p(out))
    return out


def derivative(img):
    s = _c(img, np.uint8)
    dx = np.empty(s.shape, np.int16)
    dy = np.empty(s.shape, np.int16)
    lib().orc_derivative(_p(s), s.shape[1], s.shape[0], _p(dx), _p(dy))
    return dx, dy


def point_cloud(depth, k: Intr):
    d = _c(depth, np.float32)
    out = np.empty((3,) + d.shape, np.float32)
    lib().orc_point_cloud(_p(d), d.shape[1], d.shape[0], k, _p(out))
    return out


# ---------------------------------------------------------------------- steps

def icp_step(vc, nc, vp, np_, k: Intr, Rc, tc, Rpi, tp, dist=0.1, angle=None, fp64=False):
    import math
    if angle is None:
        angle = math.sin(20.0 * 3.14159254 / 180.0)
    arrs = [_c(a, np.float32) for a in (vc, nc, vp, np_)]
    H, W = arrs[0].shape[1:]
    out = np.zeros(29, np.float64)
    mask = np.zeros((H, W), np.uint8)
    dt = np.float64 if fp64 else np.float32
    cs = [_c(a, dt) for a in (Rc, tc, Rpi, tp)]
    fn = lib().orc_icp_step_f64 if fp64 else lib().orc_icp_step
    fn(*[_p(a) for a in arrs], W, H, k, *[_p(a) for a in cs], dist, angle, _p(out), _p(mask))
    return out, mask


def rgb_residual(minScale, dIdx, dIdy, lastDepth, nextDepth, lastImage, nextImage, krkinv, kt,
                 maxDepthDelta=0.07):
    a = [_c(dIdx, np.int16), _c(dIdy, np.int16), _c(lastDepth, np.float32), _c(nextDepth, np.float32),
         _c(lastImage, np.uint8), _c(nextImage, np.uint8)]
    H, W = a[0].shape
    corres = np.zeros((H, W), DATATERM)
    cnt = np.zeros(1, np.int32)
    sig = np.zeros(1, np.int32)
    K, t = _c(krkinv, np.float32), _c(kt, np.float32)
    lib().orc_rgb_residual(minScale, *[_p(x) for x in a], W, H, maxDepthDelta, _p(K), _p(t), _p(corres),
                           _p(cnt), _p(sig))
    return corres, int(cnt[0]), int(sig[0])


def rgb_step(corres, sigma, lastDepth, k: Intr, dIdx, dIdy, sobelScale=0.125):
    c = np.ascontiguousarray(corres)
    H, W = c.shape
    out = np.zeros(29, np.float64)
    ld, dx, dy = _c(lastDepth, np.float32), _c(dIdx, np.int16), _c(dIdy, np.int16)
    lib().orc_rgb_step(_p(c), W, H, sigma, _p(ld), k, _p(dx), _p(dy), sobelScale, _p(out))
    return out


def so3_step(last, nxt, H, kinv, krlr):
    a, b = _c(last, np.uint8), _c(nxt, np.uint8)
    h, w = a.shape
    out = np.zeros(11, np.float64)
    mask = np.zeros((h, w), np.uint8)
    m = [_c(x, np.float32) for x in (H, kinv, krlr)]
    lib().orc_so3_step(_p(a), _p(b), w, h, *[_p(x) for x in m], _p(out), _p(mask))
    return out, mask


def unpack_se3(sums):
    s = _c(sums, np.float64)
    A = np.zeros(36, np.float64)
    b = np.zeros(6, np.float64)
    lib().orc_unpack_se3(_p(s), _p(A), _p(b))
    return A.reshape(6, 6), b


def ldlt_solve(A, b):
    A = _c(A, np.float64)
    b = _c(b, np.float64)
    x = np.zeros(b.shape[0], np.float64)
    z = lib().orc_ldlt_solve(b.shape[0], _p(A), _p(b), _p(x))
    return x, z


def rodrigues(r):
    r = _c(r, np.float64)
    R = np.zeros(9, np.float64)
    lib().orc_rodrigues(_p(r), _p(R))
    return R.reshape(3, 3)


def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP row loops."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def hardware_concurrency() -> int:
    """The host's online processors (std::thread::hardware_concurrency)."""
    return int(lib().orc_hardware_concurrency())


def bexpf(x: float) -> float:
    return lib().orc_bexpf(x)


# -------------------------------------------------------------------- tracker

class Odometry:
    """The oracle's RGBDOdometryef (RGBDOdometryef.h:28-131)."""

    def __init__(self, width, height, cx, cy, fx, fy, distThresh=0.10, angleThresh=None, fp64=False):
        import math
        if angleThresh is None:
            angleThresh = math.sin(20.0 * 3.14159254 / 180.0)
        self.w, self.h = width, height
        self._h = lib().orc_create(width, height, cx, cy, fx, fy, distThresh, angleThresh)
        lib().orc_set_precision(self._h, int(fp64))
        # the increment of the last call (resultRt convention), which the
        # constant-velocity start repeats; identity at a sequence's start
        self.lastRt = np.eye(4)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def initICP(self, depth, depthCutoff=20.0):
        d = _c(depth, np.uint16)
        lib().orc_init_icp_depth(self._h, _p(d), depthCutoff)

    def initICPMaps(self, v4, n4, depthCutoff=20.0):
        v, n = _c(v4, np.float32), _c(n4, np.float32)
        lib().orc_init_icp_maps(self._h, _p(v), _p(n), depthCutoff)

    def initICPModel(self, v4, n4, depthCutoff, pose):
        v, n, p = _c(v4, np.float32), _c(n4, np.float32), _c(pose, np.float32)
        lib().orc_init_icp_model(self._h, _p(v), _p(n), depthCutoff, _p(p))

    def initRGB(self, rgb):
        r = _c(rgb, np.uint8)
        lib().orc_init_rgb(self._h, _p(r), r.shape[2])

    def initRGBModel(self, rgb):
        r = _c(rgb, np.uint8)
        lib().orc_init_rgb_model(self._h, _p(r), r.shape[2])

    def initFirstRGB(self, rgb):
        r = _c(rgb, np.uint8)
        lib().orc_init_first_rgb(self._h, _p(r), r.shape[2])
        self.lastRt = np.eye(4)

    def getIncrementalTransformation(self, trans, rot, rgbOnly=False, icpWeight=10.0, pyramid=True,
                                     fastOdom=False, so3=True):
        iters = (3 if fastOdom else 10, 5 if pyramid else 0, 4 if pyramid else 0)
        t, R, _ = self.run(trans, rot, iters, rgbOnly, icpWeight, so3)
        return t, R

    def run(self, trans, rot, iters=(10, 5, 4), rgbOnly=False, icpWeight=10.0, so3=False, seed=None,
            conv=(0.0, 0.0), levels=(2, 0)):
        """orc_track_run: levels hi..lo from the seeded increment with the
        convergence test conv = (m, rad).  Returns t, R and a dict with the
        increment at the end, the solves run and the last iteration's
        statistics.  A run that ends at level 0 records its increment as the
        last call's (the constant-velocity start of the next call)."""
        t = _c(trans, np.float32).copy()
        R = _c(rot, np.float32).copy()
        it = np.ascontiguousarray(iters, dtype=np.int32)
        sd = None if seed is None else _c(seed, np.float64)
        rt = np.zeros(16, np.float64)
        nd = ctypes.c_int(0)
        lib().orc_track_run(self._h, _p(t), _p(R), int(rgbOnly), icpWeight, _p(it), int(so3),
                            None if sd is None else _p(sd), float(conv[0]), float(conv[1]), int(levels[0]),
                            int(levels[1]), _p(rt), ctypes.byref(nd))
        info = dict(self.stats(), resultRt=rt.reshape(4, 4), iterations=nd.value)
        if levels[1] == 0:
            self.lastRt = rt.reshape(4, 4).copy()
        return t, R, info

    def track_policy(self, trans, rot, pred=None, t_scale=0.05, r_scale=0.05, iters=(10, 5, 4), conv=(0.0, 0.0),
                     start="previous", so3=False, judge="full", tol=0.01, rgbOnly=False, icpWeight=10.0):
        """The library's odom_track_policy restated: the start (the predicted
        motion pred = (t / t_scale, rotation vector / r_scale), else the
        previous pose or the constant-velocity extrapolation of the last
        call's increment), the schedule with the convergence test, and the
        judge of a prediction against the constant-velocity start: both are
        refined through level 2 ("coarse") or every level ("full"), and the
        prediction falls back when its refinement's ICP residual is more than
        tol above, or its inlier count more than tol below, the other's.
        Returns t, R and a dict (kept, iterations, alt_iterations)."""
        cv = self.lastRt.copy()
        if pred is None:
            seed = cv if start == "velocity" else None
            t, R, info = self.run(trans, rot, iters, rgbOnly, icpWeight, so3 and seed is None, seed, conv)
            return t, R, {"kept": None, "iterations": info["iterations"], "alt_iterations": 0}
        q = np.asarray(pred, np.float64).reshape(6)
        M = np.eye(4)
        M[:3, :3] = rodrigues(q[3:] * r_scale)
        M[:3, 3] = q[:3] * t_scale
        # the seed is the inverse motion in the resultRt convention, built as
        # the device builds it: [R^T | -R^T t] of the float prediction
        pr = np.asarray(pred, np.float32).astype(np.float64).reshape(6)
        Rm = rodrigues(pr[3:] * np.float64(np.float32(r_scale)))
        tm = pr[:3] * np.float64(np.float32(t_scale))
        seed = np.eye(4)
        seed[:3, :3] = Rm.T
        seed[:3, 3] = -(Rm.T @ tm)
        if judge in (None, "none"):
            t, R, info = self.run(trans, rot, iters, rgbOnly, icpWeight, False, seed, conv)
            return t, R, {"kept": True, "iterations": info["iterations"], "alt_iterations": 0}

        def worse(a, b):
            ka = ("lastICPError", "lastICPCount") if not rgbOnly and icpWeight > 0 else ("lastRGBError", "lastRGBCount")
            ea, ca, eb, cb = (np.float32(a[ka[0]]), np.float32(a[ka[1]]), np.float32(b[ka[0]]),
                              np.float32(b[ka[1]]))
            one = np.float32(1.0)
            tl = np.float32(tol)
            return bool(ea > eb * (one + tl) or ca < cb * (one - tl))

        hi_lo = (2, 2) if judge == "coarse" and (iters[0] > 0 or iters[1] > 0) and iters[2] > 0 else (2, 0)
        tp, Rp, ip = self.run(trans, rot, iters, rgbOnly, icpWeight, False, seed, conv, hi_lo)
        tv, Rv, iv = self.run(trans, rot, iters, rgbOnly, icpWeight, False, cv, conv, hi_lo)
        fall = worse(ip, iv)
        win, lose = (iv, ip) if fall else (ip, iv)
        if hi_lo[1] == 0:
            t, R = (tv, Rv) if fall else (tp, Rp)
            self.lastRt = win["resultRt"].copy()
            return t, R, {"kept": not fall, "iterations": win["iterations"], "alt_iterations": lose["iterations"]}
        # the finer levels continue from the kept refinement's increment
        t, R, info = self.run(trans, rot, iters, rgbOnly, icpWeight, False, win["resultRt"], conv, (1, 0))
        return t, R, {"kept": not fall, "iterations": win["iterations"] + info["iterations"],
                      "alt_iterations": lose["iterations"]}

    def track_ex(self, trans, rot, iters=(10, 5, 4), rgbOnly=False, icpWeight=10.0, so3=False, seed=None):
        """Explicit schedule and optional seeded increment (4x4, resultRt
        convention: current = prev * seed^-1)."""
        t = _c(trans, np.float32).copy()
        R = _c(rot, np.float32).copy()
        it = np.ascontiguousarray(iters, dtype=np.int32)
        sd = None if seed is None else _c(seed, np.float64)
        lib().orc_track_ex(self._h, _p(t), _p(R), int(rgbOnly), icpWeight, _p(it), int(so3),
                           None if sd is None else _p(sd))
        return t, R

    def stats(self):
        s = np.zeros(7, np.float32)
        lib().orc_get_stats(self._h, _p(s))
        keys = ["lastICPError", "lastICPCount", "lastRGBError", "lastRGBCount", "lastSO3Error",
                "lastSO3Count", "so3Iterations"]
        return dict(zip(keys, s.tolist()))

    def so3_rotation(self):
        R = np.zeros(9, np.float64)
        lib().orc_get_so3_rotation(self._h, _p(R))
        return R.reshape(3, 3)

    def last_system(self):
        A = np.zeros(36, np.float64)
        b = np.zeros(6, np.float64)
        lib().orc_get_last_system(self._h, _p(A), _p(b))
        return A.reshape(6, 6), b

    def getCovariance(self):
        c = np.zeros(36, np.float64)
        if lib().orc_get_covariance(self._h, _p(c)) != 0:
            raise np.linalg.LinAlgError("lastA is singular")
        return c.reshape(6, 6)

    def loop_accept(self, covThresh, icpCountThresh, icpErrThresh):
        """Loop-closure acceptance of the last call (elastic_fusion_file.cpp:
        477-498): every diagonal entry of the covariance <= covThresh,
        lastICPCount > icpCountThresh and lastICPError < icpErrThresh.  A
        singular last system has no covariance and is not accepted."""
        try:
            cov = self.getCovariance()
        except np.linalg.LinAlgError:
            return False
        s = self.stats()
        return bool(np.all(np.diag(cov) <= covThresh) and s["lastICPCount"] > icpCountThresh
                    and s["lastICPError"] < icpErrThresh)

    def buf(self, name, level):
        h, w = self.h >> level, self.w >> level
        L = lib()
        for fn, dt, shape in ((L.orc_buf_f32, np.float32, None), (L.orc_buf_u8, np.uint8, None),
                              (L.orc_buf_s16, np.int16, None), (L.orc_buf_u16, np.uint16, None)):
            p = fn(self._h, name.encode(), level)
            if p:
                planes = 3 if name.startswith(("vmap", "nmap")) else 1
                n = planes * h * w
                a = np.ctypeslib.as_array(p, shape=(n,)).copy()
                return a.reshape((3, h, w) if planes == 3 else (h, w))
        raise KeyError(name)


# ------------------------------------------------------- splat and fill-in

def splat_predict(width, height, cx, cy, fx, fy, pose, model, depthCutoff, confThreshold, time, maxTime, timeDelta):
    """IndexMap::combinedPredict restated (splat_oracle.c): returns (image
    RGBA8, vertex float4, normal float4, time uint16), each height x width."""
    P = SplatParams(width, height, cx, cy, fx, fy, depthCutoff, confThreshold, time, maxTime, timeDelta)
    T = _c(pose, np.float32).reshape(16)
    m = _c(model, np.float32).reshape(-1, 12)
    image = np.zeros((height, width, 4), np.uint8)
    vertex = np.zeros((height, width, 4), np.float32)
    normal = np.zeros((height, width, 4), np.float32)
    tm = np.zeros((height, width), np.uint16)
    lib().orc_splat_predict(ctypes.byref(P), _p(T), _p(m), m.shape[0], _p(image), _p(vertex), _p(normal), _p(tm))
    return image, vertex, normal, tm


def splat_fill_in(width, height, cx, cy, fx, fy, image, vertex, normal, depth, rgb, passthroughGeometry=False,
                  passthroughImage=False):
    """FillIn::vertex / normal / image restated: returns (image RGBA8,
    vertex float4, normal float4)."""
    P = SplatParams(width, height, cx, cy, fx, fy, 0.0, 0.0, 0, 0, 0)
    im, v, n = _c(image, np.uint8), _c(vertex, np.float32), _c(normal, np.float32)
    d, r = _c(depth, np.uint16), _c(rgb, np.uint8)
    fi = np.zeros((height, width, 4), np.uint8)
    fv = np.zeros((height, width, 4), np.float32)
    fn = np.zeros((height, width, 4), np.float32)
    lib().orc_splat_fill_in(ctypes.byref(P), _p(im), _p(v), _p(n), _p(d), _p(r), r.shape[2], int(passthroughGeometry),
                            int(passthroughImage), _p(fi), _p(fv), _p(fn))
    return fi, fv, fn


# ------------------------------------------------------------ surfel model

def depth_metric(depth, maxD):
    d = _c(depth, np.uint16)
    out = np.empty(d.shape, np.float32)
    lib().orc_depth_metric(_p(d), d.size, maxD, _p(out))
    return out


class SurfelModel:
    """Model::initialise / fuse / clean and IndexMap::predictIndices restated
    (model_oracle.c)."""

    def __init__(self, width, height, cx, cy, fx, fy, max_surfels):
        self.w, self.h = width, height
        self._h = lib().orc_model_create(width, height, cx, cy, fx, fy, max_surfels)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_model_destroy(self._h)
            self._h = None

    def count(self):
        return lib().orc_model_count(self._h)

    def model(self):
        n = self.count()
        p = lib().orc_model_data(self._h)
        return np.ctypeslib.as_array(ctypes.cast(p, _F), shape=(n * 12,)).reshape(n, 12).copy() if n else \
            np.zeros((0, 12), np.float32)

    def fresh(self):
        n = lib().orc_model_fresh_count(self._h)
        p = lib().orc_model_fresh(self._h)
        return np.ctypeslib.as_array(ctypes.cast(p, _F), shape=(n * 12,)).reshape(n, 12).copy() if n else \
            np.zeros((0, 12), np.float32)

    def set(self, model):
        m = _c(model, np.float32).reshape(-1, 12)
        assert lib().orc_model_set(self._h, _p(m), m.shape[0]) == 0

    set_model = set

    def index_map(self):
        idx = np.zeros((self.h, self.w), np.uint32)
        vc = np.zeros((self.h, self.w, 4), np.float32)
        ct = np.zeros((self.h, self.w, 4), np.float32)
        nr = np.zeros((self.h, self.w, 4), np.float32)
        lib().orc_model_index_map(self._h, _p(idx), _p(vc), _p(ct), _p(nr))
        return idx, vc, ct, nr

    def initialise(self, rgb, raw, filtered, maxD, time, maxDepth):
        r, a, f = _c(rgb, np.uint8), _c(raw, np.uint16), _c(filtered, np.uint16)
        return lib().orc_model_initialise(self._h, _p(r), _p(a), _p(f), maxD, time, maxDepth)

    def predictIndices(self, pose, time, maxDepth, timeDelta):
        T = _c(pose, np.float32).reshape(16)
        lib().orc_model_predict_indices(self._h, _p(T), time, maxDepth, timeDelta)

    def fuse(self, pose, time, rgb, raw, filtered, maxD, maxDepth, weighting):
        T = _c(pose, np.float32).reshape(16)
        r, a, f = _c(rgb, np.uint8), _c(raw, np.uint16), _c(filtered, np.uint16)
        lib().orc_model_fuse(self._h, _p(T), time, _p(r), _p(a), _p(f), maxD, maxDepth, weighting)

    def clean(self, pose, time, confThreshold, timeDelta):
        T = _c(pose, np.float32).reshape(16)
        return lib().orc_model_clean(self._h, _p(T), time, confThreshold, timeDelta)
