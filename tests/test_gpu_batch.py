"""Batched tracking (odom_track_batch): pairs of uneven size in the same
launches give each pair the result it gets alone, and the oracle's.

Alone and batched differ only in how a pair's pixels are split into blocks,
i.e. in the order of FP32 block sums, so poses agree to 2e-5 m / 2e-3 deg;
against the FP32 oracle the bar is the usual 1 mm / 0.1 deg.  Graph replays
must equal the eager batch bit for bit.
"""
import numpy as np
import pytest

from slam_amd import synth
from slam_amd.odometry import Batch, OdomError, RGBDOdometry

from helpers import CUTOFF, MAXD, feed, make_pair, pose_of

pytestmark = pytest.mark.gpu

CAMS = [(synth.CAM_320, 21), (synth.CAM_640, 22), (synth.CAM_848, 23), (synth.CAM_320, 24), (synth.CAM_640, 25)]
MODES = {"joint": dict(rgbOnly=False, icpWeight=10.0, so3=True),
         "icp_only": dict(rgbOnly=False, icpWeight=100.0, so3=False),
         "rgb_only": dict(rgbOnly=True, icpWeight=10.0, so3=True)}


def mk(cam):
    return RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)


@pytest.fixture(scope="module")
def pairs(orc):
    out = []
    for cam, seed in CAMS:
        T0, T1, f0, f1 = make_pair(cam, seed, max_trans=0.03, max_rot_deg=2.0)
        out.append(dict(cam=cam, T0=T0, T1=T1, f0=f0, f1=f1, filt=orc.bilateral(f1.depth, MAXD)))
    return out


def track_alone(p, mode):
    odo = mk(p["cam"])
    feed(odo, p["f0"], p["f1"], p["T0"], p["filt"])
    t, R = odo.getIncrementalTransformation(p["T0"][:3, 3], p["T0"][:3, :3], pyramid=True, fastOdom=False,
                                            **MODES[mode])
    return pose_of(t, R), odo.stats()


def batch_handles(pairs):
    odos = [mk(p["cam"]) for p in pairs]
    for o, p in zip(odos, pairs):
        feed(o, p["f0"], p["f1"], p["T0"], p["filt"])
        o.set_pose(p["T0"][:3, 3], p["T0"][:3, :3])
    return odos


def poses_of(odos):
    out = []
    for o in odos:
        t, R = o.get_pose()
        out.append(pose_of(t, R))
    return out


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_matches_alone_and_oracle(pairs, orc, mode):
    odos = batch_handles(pairs)
    b = Batch(odos)
    m = MODES[mode]
    b.track(m["rgbOnly"], m["icpWeight"], True, False, m["so3"])
    got = poses_of(odos)
    for i, p in enumerate(pairs):
        alone, st_alone = track_alone(p, mode)
        dt, dr = synth.pose_error(got[i], alone)
        assert dt <= 2e-5 and dr <= 2e-3, (mode, i, dt, dr)
        st = odos[i].stats()
        assert st["gnIterations"] == st_alone["gnIterations"] and st["so3Iterations"] == st_alone["so3Iterations"]
        o = orc.Odometry(p["cam"].width, p["cam"].height, p["cam"].cx, p["cam"].cy, p["cam"].fx, p["cam"].fy)
        feed(o, p["f0"], p["f1"], p["T0"], p["filt"])
        t, R = o.getIncrementalTransformation(p["T0"][:3, 3], p["T0"][:3, :3], pyramid=True, fastOdom=False, **m)
        dt, dr = synth.pose_error(got[i], pose_of(t, R))
        assert dt <= 1e-3 and dr <= 0.1, (mode, i, dt, dr)


def test_batch_graph_replays_equal_eager(pairs):
    eager_odos = batch_handles(pairs)
    eager = Batch(eager_odos)
    eager.track()
    ref = poses_of(eager_odos)
    odos = batch_handles(pairs)
    b = Batch(odos)
    b.set_graph(True)
    # the same call replayed: inputs re-fed (eager work) between replays; the
    # image swap alternates the parity, so replays 1/3 and 2/4 use two graphs
    for rep in range(4):
        for o, p in zip(odos, pairs):
            if rep > 0:
                feed(o, p["f0"], p["f1"], p["T0"], p["filt"], first=True)
            o.set_pose(p["T0"][:3, 3], p["T0"][:3, :3])
        b.track()
        b.synchronize()
        got = poses_of(odos)
        if rep % 2 == 0:
            for i in range(len(pairs)):
                assert np.array_equal(got[i], ref[i]), (rep, i)
        else:
            for i in range(len(pairs)):
                dt, dr = synth.pose_error(got[i], ref[i])
                assert dt <= 2e-5 and dr <= 2e-3, (rep, i, dt, dr)


def test_batch_block_split(pairs):
    odos = batch_handles(pairs)
    b = Batch(odos)
    b.set_blocks(len(pairs))          # about one block per pair (at least one each)
    assert len(pairs) <= b.get_blocks(0, 0) <= 2 * len(pairs)
    b.track()
    few = poses_of(odos)
    odos2 = batch_handles(pairs)
    b2 = Batch(odos2)
    b2.set_blocks(4096)               # capped at 256 per pair
    assert b2.get_blocks(0, 0) <= 256 * len(pairs)
    b2.track()
    many = poses_of(odos2)
    for i in range(len(pairs)):
        dt, dr = synth.pose_error(few[i], many[i])
        assert dt <= 2e-5 and dr <= 2e-3, (i, dt, dr)


def test_batch_sequence_alternates_descriptors(orc):
    """Three frames per pair in a row with so3 (image swap after every call):
    the batch must keep matching handles tracked alone."""
    cams = [synth.CAM_320, synth.CAM_640]
    scenes = [synth.Scene(31 + i) for i in range(2)]
    trajs = [synth.spline_trajectory(4, seed=40 + i, max_step_m=0.01, max_step_deg=0.5) for i in range(2)]
    frames = [[scenes[i].render(T, cams[i], noise_seed=100 * i + k) for k, T in enumerate(trajs[i])]
              for i in range(2)]
    batch_odos = [mk(c) for c in cams]
    alone_odos = [mk(c) for c in cams]
    b = Batch(batch_odos)
    for i in range(2):
        for o in (batch_odos[i], alone_odos[i]):
            o.set_pose(trajs[i][0][:3, 3], trajs[i][0][:3, :3])
            o.initFirstRGB(frames[i][0].rgb)
    for k in range(1, 4):
        for i in range(2):
            for o in (batch_odos[i], alone_odos[i]):
                t, R = o.get_pose()
                o.initICPModel(frames[i][k - 1].vert4, frames[i][k - 1].norm4, CUTOFF,
                               pose_of(t, R).astype(np.float32))
                o.initRGBModel(frames[i][k - 1].rgba)
                o.initICP(orc.bilateral(frames[i][k].depth, MAXD), CUTOFF)
                o.initRGB(frames[i][k].rgb)
        b.track()
        for i in range(2):
            t, R = alone_odos[i].get_pose()
            alone_odos[i].getIncrementalTransformation(t, R)
        got = poses_of(batch_odos)
        want = poses_of(alone_odos)
        for i in range(2):
            dt, dr = synth.pose_error(got[i], want[i])
            assert dt <= 2e-5 and dr <= 2e-3, (k, i, dt, dr)


def test_batch_rejects_bad_sets(pairs):
    a = mk(synth.CAM_320)
    with pytest.raises(OdomError):
        Batch([a, a])
    c = RGBDOdometry(320, 240, 159.5, 119.5, 240.0, 240.0, distThresh=0.2)
    with pytest.raises(OdomError):
        Batch([a, c])
    b = Batch([a])
    with pytest.raises(OdomError) as e:
        b.track()
    assert e.value.code == -3


def test_batch_of_64_pairs_against_the_fp64_oracle(orc):
    """The benchmarked batch: 64 pairs (16 at 320x240, 32 at 640x480, 16 at
    848x480), each a seeded room and a motion of up to 3 cm / 2 deg, tracked
    by one batched ICP+RGB+SO3 call (10/5/4) with the library's batch launch
    shapes, replayed as a hipGraph.  Every pair is compared with the all-FP64
    oracle (1 mm / 0.1 deg) and with the ground truth; one pair of each size
    with the same pair tracked alone (block split only: 2e-5 m / 2e-3 deg)."""
    layout = ((synth.CAM_320, 16), (synth.CAM_640, 32), (synth.CAM_848, 16))
    cams, frames = [], []
    for cam, count in layout:
        for _ in range(count):
            seed = 500 + len(cams)
            T0, T1, f0, f1 = make_pair(cam, seed, max_trans=0.03, max_rot_deg=2.0)
            cams.append(cam)
            frames.append((T0, T1, f0, f1, orc.bilateral(f1.depth, MAXD)))
    odos = [mk(c) for c in cams]
    for o, (T0, T1, f0, f1, filt) in zip(odos, frames):
        feed(o, f0, f1, T0, filt)
        o.set_pose(T0[:3, 3], T0[:3, :3])
    b = Batch(odos)
    b.set_graph(True)
    b.track(False, 10.0, True, False, True)
    b.synchronize()
    got = poses_of(odos)
    worst = [0.0, 0.0]
    for j, (c, (T0, T1, f0, f1, filt)) in enumerate(zip(cams, frames)):
        o = orc.Odometry(c.width, c.height, c.cx, c.cy, c.fx, c.fy, fp64=True)
        feed(o, f0, f1, T0, filt)
        t, R = o.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 10.0, True, False, True)
        dt, dr = synth.pose_error(got[j], pose_of(t, R))
        worst = [max(worst[0], dt), max(worst[1], dr)]
        assert dt <= 1e-3 and dr <= 0.1, (j, c.width, dt, dr)
        st = odos[j].stats()
        assert st["gnIterations"] == 19 and not st["degenerate"], (j, st)
        assert synth.pose_error(got[j], T1)[0] < 0.01, j
    print(f"64 pairs against the FP64 oracle: worst {1e3 * worst[0]:.4f} mm, {worst[1]:.5f} deg", flush=True)
    for j in (0, 16, 48):
        T0, T1, f0, f1, filt = frames[j]
        alone, _ = track_alone(dict(cam=cams[j], T0=T0, f0=f0, f1=f1, filt=filt), "joint")
        dt, dr = synth.pose_error(got[j], alone)
        assert dt <= 2e-5 and dr <= 2e-3, (j, dt, dr)
    b.close()
