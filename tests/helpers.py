"""Shared inputs of the parity tests: seeded rendered frame pairs."""
from __future__ import annotations

import math

import numpy as np

from slam_amd import synth

ANGLE = math.sin(20.0 * 3.14159254 / 180.0)
MAXD = 5.0          # bilateral maxD (depths in the synthetic rooms reach 5 m)
CUTOFF = 20.0       # maxDepthProcessed (elastic_fusion_file.cpp:370)


def make_pair(cam, seed, max_trans=0.02, max_rot_deg=1.0, nbox=20):
    scene = synth.Scene(seed, nbox)
    rng = np.random.default_rng(seed)
    T0 = synth.base_pose()
    T1 = synth.perturb(T0, rng, max_trans, max_rot_deg)
    f0 = scene.render(T0, cam, noise_seed=2 * seed + 1)
    f1 = scene.render(T1, cam, noise_seed=2 * seed + 2)
    return T0, T1, f0, f1


def feed(odo, f0, f1, T0, filtered, first=True):
    """The reference's frame-to-model call order (elastic_fusion_file.cpp:355-371)."""
    if first:
        odo.initFirstRGB(f0.rgb)
    odo.initICPModel(f0.vert4, f0.norm4, CUTOFF, T0.astype(np.float32))
    odo.initRGBModel(f0.rgba)
    odo.initICP(filtered, CUTOFF)
    odo.initRGB(f1.rgb)


def pose_of(t, R):
    return synth.make_pose(np.asarray(R, np.float64).reshape(3, 3), np.asarray(t, np.float64).reshape(3))
