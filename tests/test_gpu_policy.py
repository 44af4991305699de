"""Tracking policies (odom_track_policy): the start (previous pose, constant
velocity, a predicted motion), the convergence test and the judge of a
prediction against the constant-velocity start -- against the oracle's
restatement (oracle.Odometry.track_policy) and against the library's own
calls that each policy reduces to (bit for bit where the launches are the
same)."""
import math

import numpy as np
import pytest

from slam_amd import synth
from slam_amd.odometry import (JUDGE_COARSE, JUDGE_FULL, JUDGE_NONE, START_PREVIOUS, START_VELOCITY, Batch, Policy,
                               RGBDOdometry)

from helpers import CUTOFF, MAXD, pose_of

pytestmark = pytest.mark.gpu

T_SCALE = R_SCALE = 0.05


def sequence(cam, seed, n=4, step_m=0.02, step_deg=1.0):
    scene = synth.Scene(seed)
    poses = synth.spline_trajectory(n, seed=seed, max_step_m=step_m, max_step_deg=step_deg)
    frames = [scene.render(T, cam, noise_seed=seed * 100 + k) for k, T in enumerate(poses)]
    return poses, frames


def mk(cam):
    return RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)


def feed_frame(odo, frames, poses, k, filt):
    """initICPModel / initRGBModel from frame k-1's rendered maps at its true
    pose, initICP / initRGB from frame k (elastic_fusion_file.cpp:366-371)."""
    odo.initICPModel(frames[k - 1].vert4, frames[k - 1].norm4, CUTOFF, poses[k - 1].astype(np.float32))
    odo.initRGBModel(frames[k - 1].rgba)
    odo.initICP(filt, CUTOFF)
    odo.initRGB(frames[k].rgb)


def motion_pred(Ta, Tb, dt=(0.0, 0.0, 0.0), dw=(0.0, 0.0, 0.0)):
    """The network's output for the camera motion Ta -> Tb (plus an error)."""
    M = np.linalg.inv(Ta) @ Tb
    w = synth.so3_log(M[:3, :3]) + np.asarray(dw)
    t = M[:3, 3] + np.asarray(dt)
    return np.concatenate([t / T_SCALE, w / R_SCALE]).astype(np.float32)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def device_pose(odo):
    t, R = odo.get_pose()
    return pose_of(t, R)


def run_device(cam, poses, frames, filts, calls):
    """Track frames 1.. with the given per-frame call (a function of the
    handle and the frame index); returns the poses and the handle."""
    odo = mk(cam)
    T0 = poses[0]
    odo.set_pose(T0[:3, 3], T0[:3, :3])
    odo.initFirstRGB(frames[0].rgb)
    out = []
    for k in range(1, len(frames)):
        feed_frame(odo, frames, poses, k, filts[k])
        calls(odo, k)
        out.append(device_pose(odo))
    return out, odo


@pytest.fixture(scope="module")
def seq(orc):
    cam = synth.CAM_640
    poses, frames = sequence(cam, 31)
    filts = [orc.bilateral(f.depth, MAXD) for f in frames]
    return cam, poses, frames, filts


def test_default_policy_is_the_reference_call(seq):
    cam, poses, frames, filts = seq
    a, oa = run_device(cam, poses, frames, filts, lambda o, k: o.track_async(False, 10.0, True, False, True))
    b, ob = run_device(cam, poses, frames, filts, lambda o, k: o.track_policy(None, Policy.make(so3=True)))
    for k, (Ta, Tb) in enumerate(zip(a, b)):
        assert np.array_equal(Ta, Tb), k
    sa, sb = oa.stats(), ob.stats()
    for key in ("lastICPError", "lastICPCount", "lastRGBError", "gnIterations", "so3Iterations"):
        assert sa[key] == sb[key], key


def oracle_sequence(orc, cam, poses, frames, filts, call):
    o = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    o.initFirstRGB(frames[0].rgb)
    T = poses[0].copy()
    out, infos = [], []
    for k in range(1, len(frames)):
        o.initICPModel(frames[k - 1].vert4, frames[k - 1].norm4, CUTOFF, poses[k - 1].astype(np.float32))
        o.initRGBModel(frames[k - 1].rgba)
        o.initICP(filts[k], CUTOFF)
        o.initRGB(frames[k].rgb)
        t, R, info = call(o, T, k)
        T = pose_of(t, R)
        out.append(T)
        infos.append(info)
    return out, infos


@pytest.mark.parametrize("start", [START_PREVIOUS, START_VELOCITY], ids=["previous", "velocity"])
def test_convergence_and_start_match_the_oracle(seq, orc, start):
    """The convergence test ends levels early (fewer than 19 solves per call)
    and the constant-velocity start repeats the last call's increment; the
    device and the oracle agree on the poses and, within one solve per call,
    on the iterations."""
    cam, poses, frames, filts = seq
    conv = (2e-4, 2e-4)
    pol = Policy.make((10, 5, 4), conv[0], conv[1], start=start, so3=False)
    its = []

    def call(o, k):
        o.track_policy(None, pol)
        its.append(o.stats()["gnIterations"])

    got, _ = run_device(cam, poses, frames, filts, call)
    name = "velocity" if start == START_VELOCITY else "previous"
    ref, infos = oracle_sequence(orc, cam, poses, frames, filts,
                                 lambda o, T, k: o.track_policy(T[:3, 3], T[:3, :3], iters=(10, 5, 4), conv=conv,
                                                                start=name))
    for k, (Tg, Tr) in enumerate(zip(got, ref)):
        dt, dr = synth.pose_error(Tg, Tr)
        assert dt <= 1e-3 and dr <= 0.1, (k, dt, dr)
    assert all(abs(a - b["iterations"]) <= 1 for a, b in zip(its, infos)), (its, infos)
    assert sum(its) < 19 * len(its)
    # tracking stays on the truth
    for k, Tg in enumerate(got):
        assert synth.pose_error(Tg, poses[k + 1])[0] < 0.01


@pytest.mark.parametrize("judge", [JUDGE_FULL, JUDGE_COARSE], ids=["full", "coarse"])
def test_judge_keeps_a_good_prediction_and_drops_a_bad_one(seq, judge):
    """A prediction 8 cm / 6 deg off always falls back, and the call equals
    the constant-velocity call; a prediction 2 mm off the true motion is
    kept (the full judge keeps it on every frame; after the coarse levels
    both refinements have converged to the same pose, and the residuals then
    decide within noise either way), and a kept call equals the unjudged call
    from it -- bit for bit, since a judged call's refinements run the same
    launches as those calls."""
    import torch
    cam, poses, frames, filts = seq
    pol = Policy.make((10, 5, 4), 1e-4, 1e-4, judge=judge)
    good = [dev(motion_pred(poses[k - 1], poses[k], dt=(0.002, 0, 0))) for k in range(len(poses))[1:]]
    bad = [dev(motion_pred(poses[k - 1], poses[k], dt=(0.08, 0, 0), dw=(0, math.radians(6), 0)))
           for k in range(len(poses))[1:]]
    torch.cuda.synchronize()

    def judged(preds, kept):
        def call(o, k):
            o.track_policy(preds[k - 1], pol, t_scale=T_SCALE, r_scale=R_SCALE)
            kept.append(o.stats()["seedUsed"])
        return call

    kept_good, kept_bad = [], []
    g, _ = run_device(cam, poses, frames, filts, judged(good, kept_good))
    b, _ = run_device(cam, poses, frames, filts, judged(bad, kept_bad))
    unjudged = Policy.make((10, 5, 4), 1e-4, 1e-4, judge=JUDGE_NONE)
    velocity = Policy.make((10, 5, 4), 1e-4, 1e-4, start=START_VELOCITY)
    g_ref, _ = run_device(cam, poses, frames, filts,
                          lambda o, k: o.track_policy(good[k - 1], unjudged, t_scale=T_SCALE, r_scale=R_SCALE))
    b_ref, _ = run_device(cam, poses, frames, filts, lambda o, k: o.track_policy(None, velocity))
    assert not any(kept_bad), kept_bad
    assert all(kept_good) if judge == JUDGE_FULL else any(kept_good), kept_good
    for k in range(len(g)):
        assert np.array_equal(b[k], b_ref[k]), k
        if all(kept_good[:k + 1]):   # the same history so far
            assert np.array_equal(g[k], g_ref[k]), k
        assert synth.pose_error(g[k], poses[k + 1])[0] < 0.01, k


@pytest.mark.parametrize("judge", ["full", "coarse"])
def test_judged_calls_match_the_oracle(seq, orc, judge):
    """A prediction a few millimetres and tenths of a degree off the truth
    (what the network gives), judged against the constant-velocity start:
    the device and the oracle keep or drop it alike and agree on the pose."""
    import torch
    cam, poses, frames, filts = seq
    rng = np.random.default_rng(5)
    raw = [motion_pred(poses[k - 1], poses[k], dt=rng.normal(0, 0.004, 3), dw=rng.normal(0, 0.004, 3))
           for k in range(1, len(poses))]
    preds = [dev(p) for p in raw]
    torch.cuda.synchronize()
    jd = JUDGE_FULL if judge == "full" else JUDGE_COARSE
    pol = Policy.make((10, 5, 4), 1e-4, 1e-4, judge=jd)
    kept = []

    def call(o, k):
        o.track_policy(preds[k - 1], pol, t_scale=T_SCALE, r_scale=R_SCALE)
        kept.append(bool(o.stats()["seedUsed"]))

    got, _ = run_device(cam, poses, frames, filts, call)
    ref, infos = oracle_sequence(orc, cam, poses, frames, filts,
                                 lambda o, T, k: o.track_policy(T[:3, 3], T[:3, :3], raw[k - 1], T_SCALE, R_SCALE,
                                                                conv=(1e-4, 1e-4), judge=judge))
    for k, (Tg, Tr) in enumerate(zip(got, ref)):
        dt, dr = synth.pose_error(Tg, Tr)
        assert dt <= 1e-3 and dr <= 0.1, (k, dt, dr)
    assert kept == [i["kept"] for i in infos]


def test_batched_policy_equals_single_pairs(orc):
    """A batch with single-call block counts runs each pair's launches of the
    single-pair call: judged predicted calls over three pairs equal the same
    calls made per handle, bit for bit."""
    import torch
    cam = synth.CAM_640
    seqs = []
    for seed in (41, 42, 43):
        poses, frames = sequence(cam, seed, n=3)
        seqs.append((poses, frames, [orc.bilateral(f.depth, MAXD) for f in frames]))
    pol = Policy.make((10, 5, 4), 1e-4, 1e-4, judge=JUDGE_COARSE)
    preds = [dev(motion_pred(p[0], p[1], dt=(0.003, -0.002, 0.0))) for p, _, _ in seqs]
    torch.cuda.synchronize()

    def prepared():
        odos = []
        for poses, frames, filts in seqs:
            o = mk(cam)
            o.set_pose(poses[0][:3, 3], poses[0][:3, :3])
            o.initFirstRGB(frames[0].rgb)
            feed_frame(o, frames, poses, 1, filts[1])
            odos.append(o)
        return odos

    single = prepared()
    for o, p in zip(single, preds):
        o.track_policy(p, pol, t_scale=T_SCALE, r_scale=R_SCALE)
    odos = prepared()
    b = Batch(odos)
    b.set_blocks(0)
    for k in range(2):
        for l in range(3):
            b.set_step_shape(k, l, *odos[0].get_step_shape(k, l))
    b.set_split_solve(False)
    pred = torch.stack(preds)
    b.track_policy(pred, pol, t_scale=T_SCALE, r_scale=R_SCALE)
    b.synchronize()
    for i, (o, s) in enumerate(zip(odos, single)):
        assert np.array_equal(device_pose(o), device_pose(s)), i
        assert o.stats()["seedUsed"] == s.stats()["seedUsed"], i
    b.close()
