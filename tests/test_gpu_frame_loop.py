"""The device frame loop (slam_amd/frame_loop.py): bilateral filter,
frame-to-model tracking against the filled-in splat prediction, fusion and
cleaning, next prediction -- against the same loop run with the CPU oracle's
components, and against the generated ground truth over a longer sequence."""
import numpy as np
import pytest

from slam_amd import evaluate, synth
from slam_amd.frame_loop import FrameLoop, LoopParams

pytestmark = pytest.mark.gpu


def frames_of(cam, seed, n, **kw):
    scene = synth.Scene(seed)
    poses = synth.spline_trajectory(n, seed=seed, **kw)
    T0i = np.linalg.inv(poses[0])
    out = []
    for k, T in enumerate(poses):
        f = scene.render(T, cam, noise_seed=seed * 1000 + k, model=False)
        out.append((T0i @ T, f))
    return out


def oracle_loop(orc, cam, frames, p):
    """FrameLoop.process with the oracle's components."""
    args = (cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    odo = orc.Odometry(*args)
    model = orc.SurfelModel(*args, p.max_surfels)
    pose = np.eye(4, dtype=np.float32)
    last_err = 0.0
    fill = None
    out = []
    for k, (_, f) in enumerate(frames):
        tick = k + 1
        filt = orc.bilateral(f.depth, p.depth_cutoff)
        if tick == 1:
            model.initialise(f.rgb, f.depth, filt, p.depth_cutoff, tick, p.max_depth)
            odo.initFirstRGB(f.rgb)
        else:
            last = pose.copy()
            ok = last_err < 1e-4
            odo.initICPModel(fill[1], fill[2], p.max_depth, pose)
            odo.initRGBModel(fill[0])
            odo.initICP(filt, p.max_depth)
            odo.initRGB(f.rgb)
            t, R = odo.getIncrementalTransformation(pose[:3, 3], pose[:3, :3], False, p.icp_weight, True, False, p.so3)
            pose = np.eye(4, dtype=np.float32)
            pose[:3, :3] = np.asarray(R, np.float32).reshape(3, 3)
            pose[:3, 3] = np.asarray(t, np.float32).reshape(3)
            last_err = odo.stats()["lastICPError"]
            diff = np.linalg.inv(pose.astype(np.float64)) @ last.astype(np.float64)
            c = np.clip((np.trace(diff[:3, :3]) - 1) / 2, -1, 1)
            w = max(float(np.linalg.norm(diff[:3, 3])), float(np.arccos(c)))
            w = max(1.0 - min(w, 0.01) / 0.01, 0.5)
            if ok:
                model.predictIndices(pose, tick, p.max_depth, p.time_delta)
                model.fuse(pose, tick, f.rgb, f.depth, filt, p.depth_cutoff, p.max_depth, w)
                model.predictIndices(pose, tick, p.max_depth, p.time_delta)
                model.clean(pose, tick, p.confidence, p.time_delta)
        img, v, n, _ = orc.splat_predict(*args, pose, model.model(), p.max_depth, p.confidence, tick, tick,
                                         p.time_delta)
        fill = orc.splat_fill_in(*args, img, v, n, filt, f.rgb)
        out.append((pose.copy(), model.count()))
    return out


def run_gpu(cam, frames, p):
    import torch
    loop = FrameLoop(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, p)
    out = []
    for _, f in frames:
        rgb, depth = torch.from_numpy(f.rgb).cuda(), torch.from_numpy(f.depth).cuda()
        pose = loop.process(rgb, depth)
        out.append((pose.copy(), loop.model.lastCount()))
    run_gpu.last = loop
    return out


def test_frame_loop_matches_the_oracle_loop(orc):
    cam = synth.CAM_320
    p = LoopParams(max_surfels=600_000)
    frames = frames_of(cam, 71, 5)
    gpu = run_gpu(cam, frames, p)
    ref = oracle_loop(orc, cam, frames, p)
    for k, ((Tg, ng), (Tr, nr)) in enumerate(zip(gpu, ref)):
        dt, dr = synth.pose_error(Tg.astype(np.float64), Tr.astype(np.float64))
        assert dt <= 1e-3 and dr <= 0.1, (k, dt, dr)
        # the models follow the poses: equal at the first frame, close after
        assert abs(ng - nr) <= 0.01 * nr + 10, (k, ng, nr)
    assert gpu[0][1] == ref[0][1]


def test_frame_loop_tracks_a_sequence():
    cam = synth.CAM_320
    frames = frames_of(cam, 72, 40, max_step_m=0.01, max_step_deg=0.5)
    gpu = run_gpu(cam, frames, LoopParams(max_surfels=1_000_000))
    stamps = [float(k) for k in range(len(frames))]
    gt = evaluate.trajectory(stamps, [T for T, _ in frames])
    est = evaluate.trajectory(stamps, [T.astype(np.float64) for T, _ in gpu])
    ate = evaluate.ate(gt, est)
    # under 1 cm of drift over the 40 frames (a lost track is tens of cm)
    assert ate["rmse"] < 0.01, ate["rmse"]
    # the model holds the room within the 3 m depth window, and repeated
    # observation has begun to make it stable (confidence above 10: about 25
    # associations at 0.35-0.5 each, a quarter of the pixels associating per
    # frame)
    M = run_gpu.last.model.model()
    assert M.shape[0] > 0.1 * cam.width * cam.height
    assert (M[:, 3] > 10.0).sum() > 100


def test_frame_loop_runs_are_bit_identical():
    """With the library's table of launch shapes (no autotuning) a run does
    not depend on timing: two runs of the same frames give the same poses
    and the same model, bit for bit."""
    cam = synth.CAM_320
    frames = frames_of(cam, 73, 12, max_step_m=0.01, max_step_deg=0.5)
    a = run_gpu(cam, frames, LoopParams(max_surfels=600_000))
    ma = run_gpu.last.model.model()
    b = run_gpu(cam, frames, LoopParams(max_surfels=600_000))
    mb = run_gpu.last.model.model()
    for k, ((Ta, na), (Tb, nb)) in enumerate(zip(a, b)):
        assert np.array_equal(Ta, Tb) and na == nb, k
    assert np.array_equal(ma.view(np.uint8), mb.view(np.uint8))


def _weights():
    import os
    return os.path.exists(os.path.join(os.path.dirname(__file__), "..", "slam_amd", "weights", "warmstart.npz"))


def test_batched_learned_loop_equals_single_sequence_loops():
    """Eight sequences in lockstep (one batched tracking call and one network
    forward of batch 8 per frame) against each sequence run alone: the same
    poses and models, bit for bit."""
    import torch
    from slam_amd.frame_loop import BatchFrameLoop
    if not _weights():
        pytest.skip("no warm-start weights in the tree")
    cam = synth.CAM_640
    n, nf = 8, 4
    seqs = [frames_of(cam, 80 + i, nf, max_step_m=0.03, max_step_deg=1.5) for i in range(n)]
    dev = [[(torch.from_numpy(f.rgb).cuda(), torch.from_numpy(f.depth).cuda()) for _, f in s] for s in seqs]
    p = LoopParams(learned=True, max_surfels=400_000, conv_t=1e-4, conv_r=1e-4)
    batch = BatchFrameLoop(n, cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, p)
    for k in range(nf):
        batch.process([dev[i][k][0] for i in range(n)], [dev[i][k][1] for i in range(n)])
    models = [m.model() for m in batch.models]
    counters = [batch.counters(i) for i in range(n)]
    batch.close()
    for i in range(n):
        one = BatchFrameLoop(1, cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, p)
        for k in range(nf):
            one.process([dev[i][k][0]], [dev[i][k][1]])
        print(f"sequence {i}: {counters[i]}", flush=True)
        for k in range(nf):
            assert np.array_equal(batch.poses[i][k], one.poses[0][k]), (i, k)
            T = seqs[i][k][0]
            assert synth.pose_error(batch.poses[i][k].astype(np.float64), T)[0] < 0.02, (i, k)
        assert np.array_equal(models[i].view(np.uint8), one.models[0].model().view(np.uint8)), i
        assert counters[i] == one.counters(0), i
        assert counters[i]["predictions"] == nf - 1
        one.close()


def test_learned_loop_tracks_a_short_sequence():
    """The learned form (network start judged against the constant-velocity
    start, convergence test) on 640x480 frames, the network's input size."""
    if not _weights():
        pytest.skip("no warm-start weights in the tree")
    cam = synth.CAM_640
    frames = frames_of(cam, 74, 12, max_step_m=0.01, max_step_deg=0.5)
    gpu = run_gpu(cam, frames, LoopParams(learned=True, max_surfels=2_000_000, conv_t=1e-4, conv_r=1e-4))
    stamps = [float(k) for k in range(len(frames))]
    gt = evaluate.trajectory(stamps, [T for T, _ in frames])
    est = evaluate.trajectory(stamps, [T.astype(np.float64) for T, _ in gpu])
    assert evaluate.ate(gt, est)["rmse"] < 0.015
    # the model holds the part of the room inside the 3 m depth window
    assert run_gpu.last.model.lastCount() > 0.1 * cam.width * cam.height
