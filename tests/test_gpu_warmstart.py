"""The learned warm start in the tracker: the seeded call against the oracle's
seeded tracker, the keep-or-fall-back rule, the network input built from the
handle's buffers, and what the trained network gains over the plain start."""
import os

import numpy as np
import pytest

from slam_amd import synth

from helpers import CUTOFF, MAXD, feed, make_pair, pose_of

pytestmark = pytest.mark.gpu

T_SCALE = R_SCALE = 0.05


def mk(cam):
    from slam_amd.odometry import RGBDOdometry
    return RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)


def pred_of(M):
    """Scaled (t, rotvec) of a camera motion M, float32."""
    return np.concatenate([M[:3, 3] / T_SCALE, synth.so3_log(M[:3, :3]) / R_SCALE]).astype(np.float32)


def seed_of(pred):
    """resultRt = M^-1 of the motion a prediction encodes (as the library)."""
    p = pred.astype(np.float64)
    M = synth.make_pose(synth.so3_exp(p[3:] * R_SCALE), p[:3] * T_SCALE)
    return np.linalg.inv(M)


@pytest.fixture(scope="module")
def big_pair(orc):
    cam = synth.CAM_640
    T0, T1, f0, f1 = make_pair(cam, 41, max_trans=0.06, max_rot_deg=4.0)
    return dict(cam=cam, T0=T0, T1=T1, f0=f0, f1=f1, filt=orc.bilateral(f1.depth, MAXD))


@pytest.mark.parametrize("iters", [(5, 2, 2), (3, 0, 0)])
def test_seeded_track_matches_oracle(orc, big_pair, iters):
    import torch
    p = big_pair
    M = np.linalg.inv(p["T0"]) @ p["T1"]
    M = synth.perturb(M, np.random.default_rng(3), 0.004, 0.2)
    pred = pred_of(M)
    gpu = mk(p["cam"])
    feed(gpu, p["f0"], p["f1"], p["T0"], p["filt"])
    gpu.set_pose(p["T0"][:3, 3], p["T0"][:3, :3])
    gpu.track_seeded(torch.from_numpy(pred).cuda(), iters=iters, select=False)
    t, R = gpu.get_pose()
    st = gpu.stats()
    assert st["seedUsed"] == 1 and st["gnIterations"] == sum(iters)
    for fp64 in (False, True):
        o = orc.Odometry(p["cam"].width, p["cam"].height, p["cam"].cx, p["cam"].cy, p["cam"].fx, p["cam"].fy,
                         fp64=fp64)
        feed(o, p["f0"], p["f1"], p["T0"], p["filt"])
        to, Ro = o.track_ex(p["T0"][:3, 3], p["T0"][:3, :3], iters=iters, so3=False, seed=seed_of(pred))
        dt, dr = synth.pose_error(pose_of(t, R), pose_of(to, Ro))
        assert dt <= 1e-3 and dr <= 0.1, (iters, fp64, dt, dr)


def test_seed_select_keeps_good_and_drops_bad(big_pair):
    import torch
    p = big_pair
    M = np.linalg.inv(p["T0"]) @ p["T1"]
    res = {}
    for name, pred in (("good", pred_of(M)), ("bad", pred_of(synth.perturb(np.eye(4), np.random.default_rng(9),
                                                                         0.25, 12.0))),
                       ("zero", np.zeros(6, np.float32))):
        gpu = mk(p["cam"])
        feed(gpu, p["f0"], p["f1"], p["T0"], p["filt"])
        gpu.set_pose(p["T0"][:3, 3], p["T0"][:3, :3])
        gpu.track_seeded(torch.from_numpy(pred).cuda(), iters=(5, 2, 2), select=True)
        t, R = gpu.get_pose()
        res[name] = (pose_of(t, R), gpu.stats())
    assert res["good"][1]["seedUsed"] == 1
    assert res["bad"][1]["seedUsed"] == 0
    # falling back means starting exactly as a zero prediction does
    assert np.array_equal(res["bad"][0], res["zero"][0])
    # and the good start tracks closer to the truth than the fallback
    assert synth.pose_error(res["good"][0], p["T1"])[0] <= synth.pose_error(res["bad"][0], p["T1"])[0] + 1e-4


def test_warmstart_input_from_handle_buffers(big_pair):
    import torch
    from slam_amd import warmstart as ws
    p = big_pair
    gpu = mk(p["cam"])
    feed(gpu, p["f0"], p["f1"], p["T0"], p["filt"])
    x = torch.zeros((1, 60, 80, 4), dtype=torch.bfloat16, device="cuda")
    gpu.initWarmStartInput(x)
    gpu.synchronize()
    net = ws.WarmStartNet(max_batch=1)
    da = torch.from_numpy(gpu.read_buffer("last_depth", 2)[None]).cuda()
    ia = torch.from_numpy(gpu.read_buffer("last_image", 2)[None]).cuda()
    db = torch.from_numpy(gpu.read_buffer("depth", 2)[None]).cuda()
    ib = torch.from_numpy(gpu.read_buffer("next_image", 2)[None]).cuda()
    ref = net.prepare(da, ia, db, ib)
    torch.cuda.synchronize()
    assert torch.equal(x.view(torch.int16), ref.view(torch.int16))
    # sanity of the channels: depths in (0, 1], intensities in [0, 1]
    xf = x.float()
    assert float(xf[..., 0].max()) <= 1.0 and float(xf[..., 2].max()) > 0.1


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                    "slam_amd", "weights", "warmstart.npz")),
                    reason="no trained weights in slam_amd/weights")
def test_trained_network_gains_over_plain_start(orc):
    """Held-out pairs with motions up to 8 cm / 6 deg: the trained network's
    prediction is closer to the true motion than the zero-motion guess, and
    the short schedule tracked from it (with the fallback rule) ends closer
    to the truth on average than the same schedule from the previous pose."""
    import torch
    from slam_amd import warmstart as ws
    net = ws.WarmStartNet(max_batch=1)
    net.load()
    cam = synth.CAM_640
    err_pred, err_zero, err_seeded, err_plain = [], [], [], []
    for k in range(8):
        T0, T1, f0, f1 = make_pair(cam, 300 + k, max_trans=0.08, max_rot_deg=6.0)
        filt = orc.bilateral(f1.depth, MAXD)
        M = np.linalg.inv(T0) @ T1
        out = {}
        for mode in ("seeded", "plain"):
            gpu = mk(cam)
            feed(gpu, f0, f1, T0, filt)
            gpu.set_pose(T0[:3, 3], T0[:3, :3])
            x = torch.zeros((1, 60, 80, 4), dtype=torch.bfloat16, device="cuda")
            gpu.initWarmStartInput(x)
            gpu.synchronize()
            pred = net.forward(x)
            torch.cuda.synchronize()
            if mode == "plain":
                pred = torch.zeros_like(pred)
            gpu.track_seeded(pred[0].contiguous(), iters=(5, 2, 2), select=(mode == "seeded"))
            t, R = gpu.get_pose()
            out[mode] = pose_of(t, R)
            if mode == "seeded":
                Mp = ws.motion_from_pred(pred[0].cpu().numpy())
                err_pred.append(synth.pose_error(Mp, M)[0])
                err_zero.append(synth.pose_error(np.eye(4), M)[0])
        err_seeded.append(synth.pose_error(out["seeded"], T1)[0])
        err_plain.append(synth.pose_error(out["plain"], T1)[0])
    assert np.mean(err_pred) < np.mean(err_zero)
    assert np.mean(err_seeded) <= np.mean(err_plain) + 1e-4, (np.mean(err_seeded), np.mean(err_plain))
