"""Build of the native libraries, in-tree (the built .so files travel with the
source tree; nothing is installed).

    python -m slam_amd.build            # everything
    python -m slam_amd.build odom       # the HIP library only

``libslam_odom.so`` is compiled for gfx950 only.  All HIP code is built with
``-ffp-contract=off`` (the per-pixel expressions that decide correspondence
masks are evaluated with one rounding per written operation) and with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["odom_kernels.hip", "odom_host.cpp", "cnn_kernels.hip", "splat_kernels.hip", "surfel_kernels.hip",
               "frame_loop.cpp", "noalloc.hip", "measure.hip"]
HIP_HEADERS = ["odom_device.h", "odom_internal.h"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_synth() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libslam_synth.so")
    _run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp",
          os.path.join(CSRC, "synth.c"), "-o", out, "-lm"])
    return out


def build_odom() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    objdir = os.path.join(LIB_DIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", path, "-o", obj])
        objs.append(obj)
    out = os.path.join(LIB_DIR, "libslam_odom.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    """The CPU oracle (test infrastructure, ``oracle/``)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all() -> None:
    build_synth()
    build_odom()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "odom": build_odom, "synth": build_synth, "oracle": build_oracle}[w]()
