"""Loading of the in-tree native libraries.

``slam_amd/_lib/libslam_odom.so`` (HIP, gfx950) is the product; it is built by
``slam_amd.build`` and is never replaced by a Python or CPU fallback: if it is
missing, loading raises.  ``libslam_synth.so`` is the scene renderer used to
make inputs.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")

_NAMES = {
    "odom": "libslam_odom.so",
    "synth": "libslam_synth.so",
}

_cache: dict[str, ctypes.CDLL] = {}


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, _NAMES[name])


def load(name: str) -> ctypes.CDLL:
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or `python -m slam_amd.build`) first")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[name] = lib
    return lib
