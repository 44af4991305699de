"""Writes tests/golden/pair96.npz: a small seeded frame pair (96x72) and the
oracle's outputs on it.  The fixture freezes the oracle: the CPU suite checks
that the oracle still reproduces it, and the GPU suite checks the HIP path
against it.  Regenerate only on a deliberate change of the oracle:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from slam_amd import synth  # noqa: E402

CAM = synth.Camera(96, 72, 72.0, 72.0, 47.5, 35.5)
MAXD, CUTOFF = 5.0, 20.0


def inputs(seed=12):
    scene = synth.Scene(seed, 20)
    rng = np.random.default_rng(seed)
    T0 = synth.base_pose()
    T1 = synth.perturb(T0, rng, 0.02, 1.0)
    f0 = scene.render(T0, CAM, noise_seed=3)
    f1 = scene.render(T1, CAM, noise_seed=4)
    return dict(T0=T0, T1=T1, vert4=f0.vert4, norm4=f0.norm4, rgba=f0.rgba, rgb0=f0.rgb,
                depth1=f1.depth, rgb1=f1.rgb)


def expected(d):
    out = {}
    filt = orc.bilateral(d["depth1"], MAXD)
    out["bilateral"] = filt
    T0 = d["T0"]
    for name, kw in (("joint", dict(icpWeight=10.0, so3=True)), ("icp", dict(icpWeight=100.0, so3=False)),
                     ("rgbonly", dict(icpWeight=10.0, so3=True, rgbOnly=True))):
        for fp64 in (False, True):
            o = orc.Odometry(CAM.width, CAM.height, CAM.cx, CAM.cy, CAM.fx, CAM.fy, fp64=fp64)
            o.initFirstRGB(d["rgb0"])
            o.initICPModel(d["vert4"], d["norm4"], CUTOFF, T0.astype(np.float32))
            o.initRGBModel(d["rgba"])
            o.initICP(filt, CUTOFF)
            o.initRGB(d["rgb1"])
            if name == "joint" and not fp64:
                for lv in range(3):
                    out[f"vmask{lv}"] = ~np.isnan(o.buf("vmap_curr", lv)[0])
                    out[f"nmask{lv}"] = ~np.isnan(o.buf("nmap_curr", lv)[0])
                    out[f"nextImage{lv}"] = o.buf("nextImage", lv)
            t, R = o.getIncrementalTransformation(T0[:3, 3].astype(np.float32), T0[:3, :3].astype(np.float32),
                                                  **kw)
            tag = f"{name}_{'f64' if fp64 else 'f32'}"
            out[f"t_{tag}"] = t
            out[f"R_{tag}"] = R
            st = o.stats()
            out[f"stats_{tag}"] = np.array([st[k] for k in sorted(st)], np.float64)
            A, b = o.last_system()
            out[f"A_{tag}"] = A
            out[f"b_{tag}"] = b
    return out


def main():
    d = inputs()
    e = expected(d)
    path = os.path.join(HERE, "pair96.npz")
    np.savez_compressed(path, **{f"in_{k}": v for k, v in d.items()}, **{f"out_{k}": v for k, v in e.items()})
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
