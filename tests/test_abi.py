"""The C ABI without a GPU: the library loads, exports exactly what
include/slam_odom.h declares, validates its arguments on the host and refuses
to run (ODOM_ERR_NODEV) where there is no gfx950 device."""
from __future__ import annotations

import ctypes
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "slam_odom.h")
CNN_HEADER = os.path.join(ROOT, "include", "slam_cnn.h")
SPLAT_HEADER = os.path.join(ROOT, "include", "slam_splat.h")
SURFEL_HEADER = os.path.join(ROOT, "include", "slam_surfel.h")
LOOP_HEADER = os.path.join(ROOT, "include", "slam_loop.h")


def declared(path=HEADER, prefix="odom_"):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*int\s+(" + prefix + r"[a-z0-9_]+)\s*\(", text, flags=re.M))


@pytest.fixture(scope="module")
def L():
    from slam_amd import _native, odometry
    if not os.path.exists(_native.lib_path("odom")):
        from slam_amd import build
        build.build_odom()
    return odometry.lib()


def test_header_and_binding_agree():
    from slam_amd import odometry
    names = declared()
    assert len(names) >= 30
    assert names == set(odometry.EXPORTS)


def test_cnn_header_and_binding_agree():
    from slam_amd import warmstart
    assert declared(CNN_HEADER, "cnn_") == set(warmstart.EXPORTS)


def test_splat_header_and_binding_agree():
    from slam_amd import splat
    assert declared(SPLAT_HEADER, "splat_") == set(splat.EXPORTS)


def test_surfel_header_and_binding_agree():
    from slam_amd import surfels
    assert declared(SURFEL_HEADER, "surfel_") == set(surfels.EXPORTS)


def test_loop_header_and_binding_agree():
    from slam_amd import frame_loop
    assert declared(LOOP_HEADER, "efloop_") == set(frame_loop.EXPORTS)


def test_library_exports_every_declared_symbol(L):
    for name in (declared() | declared(CNN_HEADER, "cnn_") | declared(SPLAT_HEADER, "splat_")
                 | declared(SURFEL_HEADER, "surfel_") | declared(LOOP_HEADER, "efloop_")):
        assert hasattr(L, name), name
    # and they are real dynamic symbols of the shared object
    from slam_amd import _native
    lib = ctypes.CDLL(_native.lib_path("odom"))
    for name in declared():
        ctypes.cast(getattr(lib, name), ctypes.c_void_p)


def test_default_params_are_the_reference_defaults(L):
    from slam_amd import odometry
    p = odometry.Params()
    assert L.odom_default_params(ctypes.byref(p), 640, 480, 319.5, 239.5, 481.2, 481.2) == 0
    assert (p.width, p.height) == (640, 480)
    assert abs(p.dist_thresh - 0.10) < 1e-7                                 # RGBDOdometryef.h:35
    assert abs(p.angle_thresh - math.sin(20.0 * 3.14159254 / 180.0)) < 1e-7  # RGBDOdometryef.h:36
    assert L.odom_default_params(None, 640, 480, 1, 1, 1, 1) == -1


def test_no_device_is_an_error_not_a_fallback(L):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from slam_amd import odometry
    p = odometry.Params()
    L.odom_default_params(ctypes.byref(p), 64, 48, 31.5, 23.5, 50.0, 50.0)
    h = ctypes.c_void_p()
    assert L.odom_create(ctypes.byref(p), ctypes.byref(h)) == -4     # ODOM_ERR_NODEV
    assert not h.value
    with pytest.raises(RuntimeError):
        odometry.RGBDOdometry(64, 48, 31.5, 23.5, 50.0, 50.0)


def test_splat_without_device_is_an_error(L):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from slam_amd import splat
    splat.lib()
    h = ctypes.c_void_p()
    assert L.splat_create(64, 48, 31.5, 23.5, 50.0, 50.0, 0, ctypes.byref(h)) == -4   # ODOM_ERR_NODEV
    assert not h.value
    assert L.splat_create(0, 48, 31.5, 23.5, 50.0, 50.0, 0, ctypes.byref(h)) == -1    # ODOM_ERR_ARG
    assert L.splat_destroy(None) == -1
    with pytest.raises(RuntimeError):
        splat.IndexMap(64, 48, 31.5, 23.5, 50.0, 50.0)
    from slam_amd import surfels
    surfels.lib()
    assert L.surfel_create(64, 48, 31.5, 23.5, 50.0, 50.0, 1000, ctypes.byref(h)) == -4
    assert L.surfel_create(64, 48, 31.5, 23.5, 50.0, 50.0, 1 << 23, ctypes.byref(h)) == -1
    assert L.surfel_destroy(None) == -1
    from slam_amd import frame_loop
    frame_loop.lib()
    P = frame_loop._Params()
    assert L.efloop_default_params(ctypes.byref(P)) == 0
    assert (P.depth_cutoff, P.max_depth, P.confidence, P.time_delta) == (3.0, 20.0, 10.0, 200)   # ef_file.cfg
    assert L.efloop_create(64, 48, 31.5, 23.5, 50.0, 50.0, ctypes.byref(P), ctypes.byref(h)) == -4
    assert L.efloop_destroy(None) == -1


def test_null_handles_are_rejected(L):
    assert L.odom_destroy(None) == -1
    assert L.odom_synchronize(None) == -1
    assert L.odom_track_async(None, 0, 10.0, 1, 0, 1) == -1


def test_missing_library_raises(monkeypatch, tmp_path):
    from slam_amd import _native
    monkeypatch.setattr(_native, "LIB_DIR", str(tmp_path))
    monkeypatch.setattr(_native, "_cache", {})
    with pytest.raises(RuntimeError, match="not built"):
        _native.load("odom")
