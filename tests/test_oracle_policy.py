"""Pins of the oracle's restatement of the tracking policies
(oracle.Odometry.run / track_policy: the convergence test, the
constant-velocity start and the judge of a predicted start) on the CPU."""
from __future__ import annotations

import numpy as np

from helpers import MAXD, pose_of
from slam_amd import synth


def constant_velocity_frames(cam, seed, n=4):
    """A camera moving by the same rigid increment every frame."""
    scene = synth.Scene(seed, 20)
    rng = np.random.default_rng(seed)
    T0 = synth.base_pose()
    M = np.linalg.inv(T0) @ synth.perturb(T0, rng, 0.02, 1.0)
    poses = [T0]
    for _ in range(n - 1):
        poses.append(poses[-1] @ M)
    return poses, [scene.render(T, cam, noise_seed=seed * 10 + k) for k, T in enumerate(poses)], M


def feed_k(o, frames, poses, k, orc):
    o.initICPModel(frames[k - 1].vert4, frames[k - 1].norm4, 20.0, poses[k - 1].astype(np.float32))
    o.initRGBModel(frames[k - 1].rgba)
    o.initICP(orc.bilateral(frames[k].depth, MAXD), 20.0)
    o.initRGB(frames[k].rgb)


def test_run_without_convergence_is_the_reference_call(orc):
    cam = synth.CAM_320
    poses, frames, _ = constant_velocity_frames(cam, 21, 2)
    out = []
    for use_run in (False, True):
        o = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
        o.initFirstRGB(frames[0].rgb)
        feed_k(o, frames, poses, 1, orc)
        T = poses[0]
        if use_run:
            t, R, info = o.run(T[:3, 3], T[:3, :3], (10, 5, 4), False, 10.0, True)
            assert info["iterations"] == 19
        else:
            t, R = o.getIncrementalTransformation(T[:3, 3], T[:3, :3])
        out.append(pose_of(t, R))
    assert np.array_equal(out[0], out[1])


def test_convergence_test_stops_levels_early_near_the_full_schedule(orc):
    """The joint ICP+RGB updates of a 2 cm / 1 deg pair stay above 0.01 mm /
    0.01 mrad through the whole 10/5/4 schedule (such a threshold changes
    nothing, bit for bit), while a 1 mm / 1 mrad threshold ends the levels
    early, within a few millimetres of the full schedule's result."""
    cam = synth.CAM_320
    poses, frames, _ = constant_velocity_frames(cam, 22, 2)
    res = {}
    for conv in ((0.0, 0.0), (1e-5, 1e-5), (1e-3, 1e-3)):
        o = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
        o.initFirstRGB(frames[0].rgb)
        feed_k(o, frames, poses, 1, orc)
        T = poses[0]
        t, R, info = o.run(T[:3, 3], T[:3, :3], (10, 5, 4), False, 10.0, False, None, conv)
        res[conv] = (pose_of(t, R), info["iterations"])
    (Tf, nf), (Tt, nt), (Tc, nc) = res[(0.0, 0.0)], res[(1e-5, 1e-5)], res[(1e-3, 1e-3)]
    assert nf == 19 and nt == 19 and np.array_equal(Tt, Tf)
    assert nc < 15
    assert synth.pose_error(Tc, Tf)[0] < 4e-3


def test_constant_velocity_start_repeats_the_last_increment(orc):
    """With no iterations the velocity start is the previous pose moved by
    the last call's increment: on a constant-velocity camera that is within
    the first call's tracking error of the truth."""
    cam = synth.CAM_320
    poses, frames, M = constant_velocity_frames(cam, 23, 3)
    o = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    o.initFirstRGB(frames[0].rgb)
    feed_k(o, frames, poses, 1, orc)
    T = poses[0]
    t, R, _ = o.track_policy(T[:3, 3], T[:3, :3])
    T1 = pose_of(t, R)
    e1 = synth.pose_error(T1, poses[1])[0]
    lastRt = o.lastRt.copy()
    feed_k(o, frames, poses, 2, orc)
    t, R, info = o.track_policy(T1[:3, 3], T1[:3, :3], iters=(0, 0, 0), start="velocity")
    T2 = pose_of(t, R)
    exp = T1.astype(np.float32).astype(np.float64) @ np.linalg.inv(lastRt)
    assert synth.pose_error(T2, exp)[0] < 1e-6
    assert synth.pose_error(T2, poses[2])[0] < 2.5 * e1 + 1e-4
    assert info["iterations"] == 0


def test_judge_drops_a_bad_prediction_and_keeps_a_good_one(orc):
    """A prediction 8 cm / 6 deg off falls back to the constant-velocity
    start (the same result as that start alone, for both judges); the true
    motion is kept by the full judge."""
    cam = synth.CAM_320
    poses, frames, M = constant_velocity_frames(cam, 24, 3)

    def pred_of(Mo):
        w = synth.so3_log(Mo[:3, :3])
        return np.concatenate([Mo[:3, 3] / 0.05, w / 0.05]).astype(np.float32)

    bad = M @ synth.make_pose(synth.so3_exp(np.array([0.0, np.radians(6), 0.0])), np.array([0.08, 0.0, 0.0]))

    def second_call(**kw):
        o = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
        o.initFirstRGB(frames[0].rgb)
        feed_k(o, frames, poses, 1, orc)
        t, R, _ = o.track_policy(poses[0][:3, 3], poses[0][:3, :3], conv=(1e-4, 1e-4))
        T1 = pose_of(t, R)
        feed_k(o, frames, poses, 2, orc)
        t, R, info = o.track_policy(T1[:3, 3], T1[:3, :3], conv=(1e-4, 1e-4), **kw)
        return pose_of(t, R), info

    Tv, _ = second_call(start="velocity")
    for judge in ("full", "coarse"):
        Tb, ib = second_call(pred=pred_of(bad), judge=judge)
        assert ib["kept"] is False
        assert synth.pose_error(Tb, Tv)[0] < 1e-6
    Tg, ig = second_call(pred=pred_of(M), judge="full")
    assert ig["kept"] is True
    assert synth.pose_error(Tg, poses[2])[0] < 0.01
