"""odom_icp_step_noalloc, the entry point that allocates nothing: its sums
equal odom_icp_step's (the same pass with the same launch shape, over the
handle's own maps) bit for bit, on two calls in a row, the second on a
smaller input; sizes beyond its maximum are refused."""
import ctypes

import numpy as np
import pytest

from slam_amd import synth
from slam_amd.odometry import ODOM_ERR_ARG, RGBDOdometry, lib

from helpers import ANGLE, MAXD, feed, make_pair

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _noalloc(maps, w, h, intr, consts, out, stream=None):
    import torch
    ptrs = [m.data_ptr() for m in maps]
    c = [np.ascontiguousarray(x, dtype=np.float32).reshape(-1) for x in consts]
    torch.cuda.synchronize()
    return lib().odom_icp_step_noalloc(*ptrs, w, h, *[float(v) for v in intr],
                                       *[x.ctypes.data for x in c], 0.1, ANGLE, out.ctypes.data, stream)


def test_noalloc_icp_step_matches_handle_entry_point(orc):
    cam = synth.CAM_640
    T0, T1, f0, f1 = make_pair(cam, 11)
    filt = orc.bilateral(f1.depth, MAXD)
    odo = RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    feed(odo, f0, f1, T0, filt)
    T = T0.copy()
    T[:3, 3] += 0.5 * (T1[:3, 3] - T0[:3, 3])
    Rc, tc = T[:3, :3].astype(np.float32), T[:3, 3].astype(np.float32)
    Rpi = np.linalg.inv(T0[:3, :3]).astype(np.float32)
    tp = T0[:3, 3].astype(np.float32)
    # level 0 (640x480), then the smaller level 1 (320x240): the second call
    # reuses the array the first one left behind
    for level in (0, 1):
        odo.set_step_shape(0, level, 256, 2)          # the no-alloc pass's launch shape
        ref, _ = odo.icp_step(level, Rc, tc, Rpi, tp, mask=False)
        w, h = cam.width >> level, cam.height >> level
        div = float(1 << level)
        maps = [_dev(odo.read_buffer(n, level)) for n in ("vmap_curr", "nmap_curr", "vmap_g_prev", "nmap_g_prev")]
        out = np.full(29, np.nan)
        rc = _noalloc(maps, w, h, (cam.fx / div, cam.fy / div, cam.cx / div, cam.cy / div), (Rc, tc, Rpi, tp), out)
        assert rc == 0, rc
        print(f"level {level}: inliers {out[28]:.0f} of {w * h}", flush=True)
        assert out[28] > 0.3 * w * h
        assert np.array_equal(out, ref), (level, np.abs(out - ref).max())
    odo.close()


def test_noalloc_rejects_sizes_beyond_its_maximum():
    import torch
    t = torch.zeros(16, dtype=torch.float32, device="cuda")
    maps = [t] * 4
    out = np.zeros(29)
    eye, zero = np.eye(3), np.zeros(3)
    consts = (eye, zero, eye, zero)
    assert _noalloc(maps, 8193, 16, (500.0, 500.0, 4096.0, 8.0), consts, out) == ODOM_ERR_ARG
    assert _noalloc(maps, 16, 8193, (500.0, 500.0, 8.0, 4096.0), consts, out) == ODOM_ERR_ARG
    assert _noalloc(maps, 0, 16, (500.0, 500.0, 8.0, 8.0), consts, out) == ODOM_ERR_ARG
    # a tiny all-NaN input runs and has no inliers
    nan = torch.full((3 * 4 * 4,), float("nan"), dtype=torch.float32, device="cuda")
    out[:] = -1.0
    assert _noalloc([nan] * 4, 4, 4, (3.0, 3.0, 1.5, 1.5), consts, out) == 0
    assert np.all(out == 0.0)
