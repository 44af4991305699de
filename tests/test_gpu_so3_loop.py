"""The SO3 pre-alignment loop as one launch with a bounded grid barrier
(odom_set_so3_loop) against the launch-per-iteration form: the same poses
and statistics, bit for bit, alone and in a batch of one pair per size; and
the failure path of the barrier, forced from the host (a spin limit of 0
gives up at the first wait), which every workgroup leaves and which the
handle reports as ODOM_ERR_HIP until initFirstRGB."""
import numpy as np
import pytest

from slam_amd import synth
from slam_amd.odometry import Batch, OdomError, RGBDOdometry

from helpers import CUTOFF, MAXD, pose_of

pytestmark = pytest.mark.gpu


def mk(cam):
    return RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)


def sequence(cam, seed, n=5):
    scene = synth.Scene(seed)
    poses = synth.spline_trajectory(n, seed=seed, max_step_m=0.02, max_step_deg=1.5)
    frames = [scene.render(T, cam, noise_seed=seed * 100 + k) for k, T in enumerate(poses)]
    return poses, frames


def feed_frame(odo, frames, poses, k, filt):
    odo.initICPModel(frames[k - 1].vert4, frames[k - 1].norm4, CUTOFF, poses[k - 1].astype(np.float32))
    odo.initRGBModel(frames[k - 1].rgba)
    odo.initICP(filt, CUTOFF)
    odo.initRGB(frames[k].rgb)


def run(odo, poses, frames, filts):
    odo.set_pose(poses[0][:3, 3], poses[0][:3, :3])
    odo.initFirstRGB(frames[0].rgb)
    out, stats = [], []
    for k in range(1, len(frames)):
        feed_frame(odo, frames, poses, k, filts[k])
        odo.track_async(False, 10.0, True, False, True)
        t, R = odo.get_pose()
        out.append(pose_of(t, R))
        stats.append(odo.stats())
    return out, stats


@pytest.mark.parametrize("cam,seed", [(synth.CAM_640, 61), (synth.CAM_320, 62)], ids=["640x480", "320x240"])
def test_one_launch_loop_equals_launch_per_iteration(orc, cam, seed):
    poses, frames = sequence(cam, seed)
    filts = [orc.bilateral(f.depth, MAXD) for f in frames]
    loop = mk(cam)
    per_it = mk(cam)
    per_it.set_so3_loop(-1)
    a, sa = run(loop, poses, frames, filts)
    b, sb = run(per_it, poses, frames, filts)
    for k in range(len(a)):
        assert np.array_equal(a[k], b[k]), k
        for key in ("so3Iterations", "lastSO3Error", "lastSO3Count", "gnIterations", "lastICPError"):
            assert sa[k][key] == sb[k][key], (k, key)
        assert sa[k]["so3Iterations"] >= 1
        assert synth.pose_error(a[k], poses[k + 1])[0] < 0.01, k


def test_loop_blocks_compute_the_derivatives(orc):
    """The derivative images and photometric source masks that spare blocks
    of the one-launch loop compute equal the standalone derivative pass's, bit
    for bit, at every level (a ragged size, so the blocks' last segment is
    partial)."""
    cam = synth.Camera(202, 154, 150.0, 150.0, 100.5, 76.5)
    poses, frames = sequence(cam, 66, n=2)
    filt = orc.bilateral(frames[1].depth, MAXD)
    ref, loop = mk(cam), mk(cam)
    for o in (ref, loop):
        o.set_pose(poses[0][:3, 3], poses[0][:3, :3])
        o.initFirstRGB(frames[0].rgb)
        feed_frame(o, frames, poses, 1, filt)
    ref.compute_derivatives()
    want = [(ref.read_buffer("dxy", l), ref.read_buffer("zsrc", l)) for l in range(3)]
    loop.track_async(False, 10.0, True, False, True)
    loop.synchronize()
    # the tracking call swapped the image pair afterwards; the derivative
    # buffers are not swapped
    for l in range(3):
        dxy, zsrc = loop.read_buffer("dxy", l), loop.read_buffer("zsrc", l)
        assert np.array_equal(dxy, want[l][0]), l
        assert np.array_equal(zsrc.view(np.uint32), want[l][1].view(np.uint32)), l


def test_batched_one_launch_loop_equals_per_pair_calls(orc):
    """Three pairs of different sizes in one batch (48 workgroups of the
    loop, one barrier counter per pair): each pair's pose equals its own
    call's (single-call block counts, so the same launches per pair)."""
    cams = [(synth.CAM_320, 63), (synth.CAM_640, 64), (synth.CAM_848, 65)]
    seqs = []
    for cam, seed in cams:
        poses, frames = sequence(cam, seed, n=2)
        seqs.append((cam, poses, frames, orc.bilateral(frames[1].depth, MAXD)))

    def prepared():
        odos = []
        for cam, poses, frames, filt in seqs:
            o = mk(cam)
            o.set_pose(poses[0][:3, 3], poses[0][:3, :3])
            o.initFirstRGB(frames[0].rgb)
            feed_frame(o, frames, poses, 1, filt)
            odos.append(o)
        return odos

    single = prepared()
    for o in single:
        o.track_async(False, 10.0, True, False, True)
    odos = prepared()
    b = Batch(odos)
    b.set_blocks(0)
    for k in range(2):
        for l in range(3):
            b.set_step_shape(k, l, *odos[0].get_step_shape(k, l))
    b.set_split_solve(False)
    b.track(False, 10.0, True, False, True)
    b.synchronize()
    for i, (o, s) in enumerate(zip(odos, single)):
        if i == 0:
            assert np.array_equal(pose_of(*o.get_pose()), pose_of(*s.get_pose())), i
        # other sizes run with the first handle's launch shapes: the block
        # split may differ, so the poses agree to the block-sum order
        dt, dr = synth.pose_error(pose_of(*o.get_pose()), pose_of(*s.get_pose()))
        assert dt <= 2e-5 and dr <= 2e-3, (i, dt, dr)
        assert o.stats()["so3Iterations"] == s.stats()["so3Iterations"], i
    b.close()


def test_barrier_failure_is_reported_and_cleared(orc):
    cam = synth.CAM_320
    poses, frames = sequence(cam, 66, n=3)
    filts = [orc.bilateral(f.depth, MAXD) for f in frames]
    odo = mk(cam)
    odo.set_so3_loop(0)
    odo.set_pose(poses[0][:3, 3], poses[0][:3, :3])
    odo.initFirstRGB(frames[0].rgb)
    feed_frame(odo, frames, poses, 1, filts[1])
    odo.track_async(False, 10.0, True, False, True)
    with pytest.raises(OdomError) as e:
        odo.get_pose()
    assert e.value.code == -2
    with pytest.raises(OdomError):
        odo.stats()
    # the counters were reset by the last workgroup out: the loop works again
    odo.set_so3_loop()
    got, _ = run(odo, poses, frames, filts)
    ref, _ = run(mk(cam), poses, frames, filts)
    for k in range(len(got)):
        assert np.array_equal(got[k], ref[k]), k
