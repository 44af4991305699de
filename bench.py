#!/usr/bin/env python
"""Frames/s of dense RGB-D odometry on one MI355X.

Workload: a 640x480 frame-to-model sequence (seeded procedural room, Perlin
textures, depth noise 1 mm * z^2, 2 % dropout) on a smooth spline of up to
1 cm / 0.5 deg per frame.  Each step tracks one frame exactly as the
reference's frame loop does (elastic_fusion_file.cpp:342-374): bilateral
filter, initICPModel (model maps of the previous frame at the current
estimate), initRGBModel, initICP, initRGB, then getIncrementalTransformation
with ICP+RGB (icpWeight 10), 3 levels at 10/5/4 iterations and the SO(3)
pre-alignment.  Inputs are resident in HBM before timing starts; every call
is enqueued without host synchronisation and the pose chain stays on the
device.

Prints one JSON line: throughput, the roofline of the dominant step kernel
(HIP events on the library's stream), the CPU restatement's frames/s on a
bounded sample of the same sequence, and the pose error against the FP64
oracle on that sample.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MAXD = 5.0
CUTOFF = 20.0
ICP_WEIGHT = 10.0
# algorithmic bytes per pixel and iteration (unique data each pass must touch)
BYTES_ICP = 48   # v_c, n_c, v_p, n_p: 4 x 3 FP32
BYTES_RGB = 14   # I_next 1, dIdx 2, dIdy 2, z_next 4, z_last 4, I_last 1
# k_step_a computes the ICP products and the photometric association, so it
# reads the ICP bytes plus the association's z_next, I_next, z_last, I_last;
# k_step_b then reads the gradients (the rest of the RGB bytes)
BYTES_A = BYTES_ICP + 10
BYTES_B = BYTES_RGB - 10


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


BATCH_LAYOUT = (("CAM_320", 16), ("CAM_640", 32), ("CAM_848", 16))


def batched_pairs(args) -> dict:
    """64 independent pairs of uneven size (16 at 320x240, 32 at 640x480, 16
    at 848x480), seeded rooms, per-pair motion up to 3 cm / 2 deg, tracked by
    one batched call (ICP+RGB+SO3, 10/5/4) replayed as a hipGraph.  Reports
    pairs/s of the batched call (inputs prepared per handle beforehand), the
    roofline of the batched ICP pass, and the pose error of a few pairs against
    the FP64 oracle."""
    from slam_amd import synth
    from slam_amd.odometry import Batch, RGBDOdometry

    t0 = time.time()
    cams, frames = [], []
    i = 0
    for cam_name, count in BATCH_LAYOUT:
        cam = getattr(synth, cam_name)
        for _ in range(count):
            seed = 500 + i
            scene = synth.Scene(seed)
            rng = np.random.default_rng(seed)
            T0 = synth.base_pose()
            T1 = synth.perturb(T0, rng, 0.03, 2.0)
            f0 = scene.render(T0, cam, noise_seed=2 * seed + 1)
            f1 = scene.render(T1, cam, noise_seed=2 * seed + 2)
            cams.append(cam)
            frames.append((T0, T1, f0, f1))
            i += 1
    log(f"batch: rendered {len(frames)} pairs in {time.time() - t0:.1f}s")
    odos = [RGBDOdometry(c.width, c.height, c.cx, c.cy, c.fx, c.fy) for c in cams]

    def feed_all():
        for o, (T0, T1, f0, f1) in zip(odos, frames):
            o.set_pose(T0[:3, 3], T0[:3, :3])
            o.initFirstRGB(f0.rgb)
            o.initICPModel(f0.vert4, f0.norm4, CUTOFF, T0.astype(np.float32))
            o.initRGBModel(f0.rgba)
            o.initICPRaw(f1.depth, MAXD, CUTOFF)
            o.initRGB(f1.rgb)
        for o in odos:
            o.synchronize()

    b = Batch(odos)
    b.set_graph(True)
    feed_all()
    b.track(False, ICP_WEIGHT, True, False, True)
    b.synchronize()
    got = [o.get_pose() for o in odos]

    # accuracy: a few pairs of each size against the FP64 oracle
    acc = None
    if not args.no_cpu and args.batch_oracle_pairs > 0:
        from oracle import oracle as orc
        pick = []
        k = 0
        for _, count in BATCH_LAYOUT:
            pick += list(range(k, k + max(1, args.batch_oracle_pairs // len(BATCH_LAYOUT))))
            k += count
        errs = []
        for j in pick:
            T0, T1, f0, f1 = frames[j]
            c = cams[j]
            o = orc.Odometry(c.width, c.height, c.cx, c.cy, c.fx, c.fy, fp64=True)
            o.initFirstRGB(f0.rgb)
            o.initICPModel(f0.vert4, f0.norm4, CUTOFF, T0.astype(np.float32))
            o.initRGBModel(f0.rgba)
            o.initICP(orc.bilateral(f1.depth, MAXD), CUTOFF)
            o.initRGB(f1.rgb)
            t, R = o.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, ICP_WEIGHT, True, False, True)
            ref = synth.make_pose(R.astype(np.float64), t.astype(np.float64))
            g = synth.make_pose(got[j][1].astype(np.float64), got[j][0].astype(np.float64))
            errs.append(synth.pose_error(g, ref))
        acc = {"pairs": len(pick), "vs_fp64_oracle_mean_mm": round(1e3 * float(np.mean([e[0] for e in errs])), 5),
               "vs_fp64_oracle_max_mm": round(1e3 * float(np.max([e[0] for e in errs])), 5),
               "vs_fp64_oracle_max_deg": round(float(np.max([e[1] for e in errs])), 6)}
        log(f"batch accuracy {acc}")

    # launch shapes for the batch's own launches
    b.autotune(False, ICP_WEIGHT)
    bshapes = {f"{'ab'[k]}{l}": list(b.get_step_shape(k, l)) for k in range(2) for l in range(3)}
    log(f"batch shapes {bshapes}")

    # timing: repeated batched calls on the prepared pairs (each call swaps the
    # pairs' images, as the reference does, so consecutive calls alternate)
    for _ in range(3):
        b.track(False, ICP_WEIGHT, True, False, True)
    b.synchronize()
    reps = max(1, args.batch_reps)
    t0 = time.perf_counter()
    for _ in range(reps):
        b.track(False, ICP_WEIGHT, True, False, True)
    b.synchronize()
    call_s = (time.perf_counter() - t0) / reps
    prof = [b.profile_track(False, ICP_WEIGHT, True, False, True) for _ in range(5)]
    iters = [10, 5, 4]
    npx = [sum((c.width >> l) * (c.height >> l) for c in cams) for l in range(3)]
    us_a = [float(np.mean([p["us_a"][l] for p in prof])) for l in range(3)]
    us_b = [float(np.mean([p["us_b"][l] for p in prof])) for l in range(3)]
    launches = sum(iters)
    bytes_a = sum(iters[l] * BYTES_A * npx[l] for l in range(3)) / launches
    avg_a = sum(iters[l] * us_a[l] for l in range(3)) / launches
    ach = bytes_a / (avg_a * 1e-6) / 1e9
    l0_gbs = BYTES_A * npx[0] / (us_a[0] * 1e-6) / 1e9
    joint = sum(iters[l] * (BYTES_ICP + BYTES_RGB) * npx[l] for l in range(3)) / \
        (sum(iters[l] * (us_a[l] + us_b[l]) for l in range(3)) * 1e-6) / 1e9
    out = {"workload": "64 pairs: 16x320x240 + 32x640x480 + 16x848x480, <=3cm/2deg, ICP+RGB+SO3 10/5/4, "
                       "one batched call per step (hipGraph)",
           "pairs_per_s": round(len(odos) / call_s, 1), "ms_per_call": round(call_s * 1e3, 3),
           "roofline": {"bound": "hbm", "kernel": "k_step_a (batched)", "achieved": round(ach, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "level0_GBs": round(l0_gbs, 1), "bytes_per_launch": int(bytes_a),
                        "avg_launch_us": round(avg_a, 2)},
           "per_level_us": {"k_step_a": [round(x, 2) for x in us_a], "k_step_b": [round(x, 2) for x in us_b]},
           "joint_iteration_gbs": round(joint, 1),
           "blocks_per_launch": [b.get_blocks(0, l) for l in range(3)],
           "shapes": bshapes,
           "accuracy": acc}
    log(f"batch {out}")
    b.close()
    for o in odos:
        o.close()
    return out


def _splat_traffic():
    """HBM bytes per raster launch from the committed counter runs (FETCH_SIZE
    + WRITE_SIZE, copy-scaled) of the same 2M-surfel workload, or None."""
    path = os.path.join(ROOT, "profiles", "k_splat_raster_traffic_640x480_2M.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return int(json.load(fh)["traffic_bytes"])


def splat_prediction(args) -> dict:
    """The model prediction that feeds frame-to-model tracking: a 640x480
    combinedPredict of a 2M-surfel model (48 B per surfel: seven views of one
    room, in insertion order) and its fill-in, with the reference configs'
    confidence 10, timeDelta 200 and depth cutoff 20 m.  Per-launch times come
    from HIP events on the handle's stream; the raster pass's roofline counts
    the model bytes it must read.  One prediction and fill-in are checked bit
    for bit against the oracle, whose time for that prediction on one host
    core is the CPU baseline."""
    import torch
    from slam_amd import splat, synth

    cam = synth.CAM_640
    scene = synth.Scene(900)
    poses = synth.spline_trajectory(8, seed=900, max_step_m=0.12, max_step_deg=8.0)
    t0 = time.time()
    parts = []
    for k, T in enumerate(poses[:7]):
        f = scene.render(T, cam, noise_seed=900 + k, sensor=False)
        parts.append(synth.surfels_from_frame(f, cam, conf=10.0 + k, init_time=k, last_time=100 + k))
    model = np.ascontiguousarray(np.concatenate(parts)[:args.splat_surfels])
    n = model.shape[0]
    view = scene.render(poses[3], cam, noise_seed=990)
    log(f"splat: {n} surfels built in {time.time() - t0:.1f}s")
    pose = poses[3].astype(np.float32)
    P = dict(depthCutoff=20.0, confThreshold=10.0, time=110, maxTime=110, timeDelta=200)
    dev_model = torch.from_numpy(model).cuda()
    dev_depth = torch.from_numpy(view.depth).cuda()
    dev_rgb = torch.from_numpy(view.rgb).cuda()
    im = splat.IndexMap(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    im.profile(pose, dev_model, P["depthCutoff"], P["confThreshold"], P["time"], P["maxTime"], P["timeDelta"],
               dev_depth, dev_rgb, reps=5)
    prof = im.profile(pose, dev_model, P["depthCutoff"], P["confThreshold"], P["time"], P["maxTime"],
                      P["timeDelta"], dev_depth, dev_rgb, reps=args.splat_reps)
    total_us = prof["raster_us"] + prof["resolve_us"] + prof["fill_us"]
    model_bytes = 48 * n
    out_bytes = cam.width * cam.height * (8 + 8 + 38)          # keys read + cleared, outputs
    fill_bytes = cam.width * cam.height * (36 + 2 + 3 + 36)    # prediction + sensor in, fill out
    res = {
        "workload": f"640x480 combinedPredict + fill-in of a {n}-surfel model (48 B each), confidence 10, "
                    "timeDelta 200, cutoff 20 m",
        "surfels": n,
        "predictions_per_s": round(1e6 / total_us, 1),
        "us": {k: round(prof[k], 2) for k in ("raster_us", "resolve_us", "fill_us")},
        "fragments": prof["fragments"],
        "roofline": {"bound": "hbm", "kernel": "k_splat_raster", "achieved": round(model_bytes / prof["raster_us"] / 1e3, 1),
                     "peak": 8000.0, "unit": "GB/s",
                     "frac": round(model_bytes / prof["raster_us"] / 1e3 / 8000.0, 4),
                     "bytes_per_launch": model_bytes, "avg_launch_us": round(prof["raster_us"], 2),
                     "traffic": _splat_traffic()},
        "resolve_gbs": round(out_bytes / prof["resolve_us"] / 1e3, 1),
        "fill_gbs": round(fill_bytes / prof["fill_us"] / 1e3, 1) if prof["fill_us"] else None,
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        im.combinedPredict(pose, dev_model, P["depthCutoff"], P["confThreshold"], P["time"], P["maxTime"],
                           P["timeDelta"])
        im.fillIn(dev_depth, dev_rgb)
        got = [im.read(w) for w in (splat.IMAGE, splat.VERTEX, splat.NORMAL, splat.TIME)]
        fill = [im.read(w) for w in (splat.FILL_IMAGE, splat.FILL_VERTEX, splat.FILL_NORMAL)]
        t1 = time.time()
        exp = orc.splat_predict(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, pose, model,
                                P["depthCutoff"], P["confThreshold"], P["time"], P["maxTime"], P["timeDelta"])
        cpu_s = time.time() - t1
        expf = orc.splat_fill_in(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, *exp[:3], view.depth,
                                 view.rgb)
        same = all(np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
                   for a, b in zip(got + fill, list(exp) + list(expf)))
        res["bit_exact_vs_oracle"] = bool(same)
        res["cpu_baseline"] = {"value": round(1.0 / cpu_s, 3), "unit": "predictions/s", "cores": 1, "kind": "port",
                               "sample": "one prediction of the same model and pose, oracle/splat_oracle.c "
                                         "(in-order GL_LESS draw, scalar)"}
    log(f"splat {res}")
    return res


def surfel_fusion(args) -> dict:
    """The model update that follows each tracked frame: at 640x480, a model
    grown by fusing a sequence of one room (initialise, then per frame index
    map, fuse, index map, clean with confidence 10 and timeDelta 200), then
    one more frame's update timed with HIP events per stage (the same frame
    fused into the same starting model, which is restored before each
    repetition).  The update of that frame is checked bit for bit against the
    oracle, whose time for it on one host core is the CPU baseline."""
    import torch
    from slam_amd import surfels, synth
    from oracle import oracle as orc

    cam = synth.CAM_640
    scene = synth.Scene(910)
    n = args.fusion_frames + 1
    poses = synth.spline_trajectory(n, seed=910, max_step_m=0.02, max_step_deg=1.0)
    T0i = np.linalg.inv(poses[0])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    frames = []
    for k, T in enumerate(poses):
        f = scene.render(T, cam, noise_seed=9100 + k, model=False)
        frames.append(((T0i @ T).astype(np.float32), f, orc.bilateral(f.depth, 5.0)))
    args_cam = (cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    P = dict(depthCutoff=5.0, maxDepth=20.0, conf=10.0, delta=200)
    m = surfels.SurfelModel(*args_cam, 4_000_000)
    t0 = time.time()
    for k, (T, f, filt) in enumerate(frames[:-1]):
        if k == 0:
            m.initialise(dev(f.rgb), dev(f.depth), dev(filt), P["depthCutoff"], 1, P["maxDepth"])
            continue
        m.predictIndices(T, k + 1, P["maxDepth"], P["delta"])
        m.fuse(T, k + 1, dev(f.rgb), dev(f.depth), dev(filt), P["depthCutoff"], P["maxDepth"], 1.0)
        m.predictIndices(T, k + 1, P["maxDepth"], P["delta"])
        m.clean(T, k + 1, P["conf"], P["delta"])
    start = m.model()
    log(f"fusion: model of {start.shape[0]} surfels from {n - 1} frames in {time.time() - t0:.1f}s")
    T, f, filt = frames[-1]
    tick = n
    prof = m.profile(T, tick, dev(f.rgb), dev(f.depth), dev(filt), P["depthCutoff"], P["maxDepth"], 1.0, P["conf"],
                     P["delta"], reps=args.fusion_reps)
    total = sum(prof.values())
    res = {
        "workload": f"640x480 frame fused into a {start.shape[0]}-surfel model: predictIndices, fuse, "
                    "predictIndices, clean (confidence 10, timeDelta 200)",
        "surfels": int(start.shape[0]),
        "frames_per_s": round(1e6 / total, 1),
        "us": {k: round(v, 2) for k, v in prof.items()},
    }
    if not args.no_cpu:
        got = m.model()
        ref = orc.SurfelModel(*args_cam, 4_000_000)
        ref.set(start)
        t1 = time.time()
        ref.predictIndices(T, tick, P["maxDepth"], P["delta"])
        ref.fuse(T, tick, f.rgb, f.depth, filt, P["depthCutoff"], P["maxDepth"], 1.0)
        ref.predictIndices(T, tick, P["maxDepth"], P["delta"])
        ref.clean(T, tick, P["conf"], P["delta"])
        cpu_s = time.time() - t1
        exp = ref.model()
        res["bit_exact_vs_oracle"] = bool(got.shape == exp.shape and np.array_equal(got.view(np.uint8),
                                                                                   exp.view(np.uint8)))
        res["cpu_baseline"] = {"value": round(1.0 / cpu_s, 3), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": "the same frame's update of the same model, oracle/model_oracle.c (scalar)"}
    log(f"fusion {res}")
    return res


def frame_loop_sequence(args) -> dict:
    """The reference's frame loop without loop closure on the device
    (slam_amd/frame_loop.py): per frame the bilateral filter, frame-to-model
    ICP+RGB+SO3 tracking against the filled-in splat prediction of the surfel
    model, the fusion and cleaning, and the next prediction, with the
    reference configs' values (depth cutoff 3 m, confidence 10, timeDelta
    200).  A 640x480 sequence of one room with bursts up to 10 cm / 8 deg per
    frame, frames resident on the device; reports the whole loop's frames/s
    (host reads of the pose and the surfel count included, as the reference
    loop makes them), the model size, and ATE / RPE against the generated
    ground truth."""
    import torch
    from slam_amd import evaluate, synth
    from slam_amd.frame_loop import FrameLoop

    cam = synth.CAM_640
    n = args.loop_frames
    poses = synth.spline_trajectory(n, seed=920, max_step_m=0.10, max_step_deg=8.0, bursts=True)
    T0i = np.linalg.inv(poses[0])
    gt = [T0i @ T for T in poses]
    scene = synth.Scene(920)
    t0 = time.time()
    frames = []
    for k, T in enumerate(poses):
        f = scene.render(T, cam, noise_seed=92000 + k, model=False)
        frames.append((torch.from_numpy(f.rgb).cuda(), torch.from_numpy(f.depth).cuda()))
        if (k + 1) % 500 == 0:
            log(f"frame loop: rendered {k + 1}/{n} frames")
    torch.cuda.synchronize()
    log(f"frame loop: {n} frames resident in {time.time() - t0:.1f}s")
    steps = [np.linalg.inv(gt[k]) @ gt[k + 1] for k in range(n - 1)]
    from slam_amd.frame_loop import LoopParams

    def run(learned: bool) -> dict:
        loop = FrameLoop(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, LoopParams(learned=learned))
        warm = min(5, n)
        for k in range(warm):
            loop.process(*frames[k])
        torch.cuda.synchronize()
        t1 = time.time()
        for k in range(warm, n):
            loop.process(*frames[k])
        torch.cuda.synchronize()
        el = time.time() - t1
        stamps = [float(k) for k in range(n)]
        gtt = evaluate.trajectory(stamps, gt)
        est = evaluate.trajectory(stamps, [T.astype(np.float64) for T in loop.poses])
        ate = evaluate.ate(gtt, est)
        rpe = evaluate.rpe(gtt, est, delta=1, delta_unit="f")
        M = loop.model.model()
        st = loop.odo.stats()
        r = {
            "frames_per_s": round((n - warm) / el, 1),
            "ms_per_frame": round(1e3 * el / (n - warm), 3),
            "surfels_final": int(M.shape[0]),
            "surfels_stable": int((M[:, 3] > 10.0).sum()),
            "ate_rmse_mm": round(1e3 * ate["rmse"], 2),
            "ate_max_mm": round(1e3 * ate["max"], 2),
            "rpe_rmse_mm": round(1e3 * rpe["translational"]["rmse"], 3),
            "rpe_rmse_deg": round(rpe["rotational_deg"]["rmse"], 4),
            "frames_lost_gt_5cm": int(sum(np.linalg.norm((np.linalg.inv(g) @ e.astype(np.float64))[:3, 3]) > 0.05
                                          for g, e in zip(gt, loop.poses))),
        }
        lost = [k for k, (g, e) in enumerate(zip(gt, loop.poses))
                if np.linalg.norm((np.linalg.inv(g) @ e.astype(np.float64))[:3, 3]) > 0.05]
        r["first_lost_frame"] = lost[0] if lost else None
        if learned:
            r["prediction_kept_frac"] = round(st["seedKept"] / max(1, st["seedCalls"]), 3)
        return r

    # the loop's own speed on a sequence it tracks throughout (smooth, 300 frames)
    sposes = synth.spline_trajectory(300, seed=921, max_step_m=0.02, max_step_deg=1.0)
    sframes = [(torch.from_numpy(f.rgb).cuda(), torch.from_numpy(f.depth).cuda()) for f in
               (scene.render(T, cam, noise_seed=93000 + k, model=False) for k, T in enumerate(sposes))]
    torch.cuda.synchronize()
    sloop = FrameLoop(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    for k in range(5):
        sloop.process(*sframes[k])
    t2 = time.time()
    for k in range(5, len(sframes)):
        sloop.process(*sframes[k])
    torch.cuda.synchronize()
    sel = time.time() - t2
    s0i = np.linalg.inv(sposes[0])
    sgt = evaluate.trajectory([float(k) for k in range(len(sposes))], [s0i @ T for T in sposes])
    sest = evaluate.trajectory([float(k) for k in range(len(sposes))], [T.astype(np.float64) for T in sloop.poses])
    smooth = {"frames": len(sposes), "frames_per_s": round((len(sposes) - 5) / sel, 1),
              "ms_per_frame": round(1e3 * sel / (len(sposes) - 5), 3), "surfels_final": sloop.model.lastCount(),
              "ate_rmse_mm": round(1e3 * evaluate.ate(sgt, sest)["rmse"], 2),
              "workload": "300 frames on a <=2 cm / 1 deg spline, the same room, native loop"}
    sloop.close()
    del sframes
    res = {
        "smooth": smooth,
        "workload": f"{n}-frame 640x480 frame loop (track, fuse, clean, predict), bursts up to "
                    f"{1e3 * max(np.linalg.norm(s[:3, 3]) for s in steps):.0f} mm / "
                    f"{max(np.degrees(np.arccos(np.clip((np.trace(s[:3, :3]) - 1) / 2, -1, 1))) for s in steps):.1f}"
                    " deg per frame, no loop closure",
        "frames": n,
        "plain": run(False),
        "learned": run(True) if os.path.exists(os.path.join(ROOT, "slam_amd", "weights", "warmstart.npz")) else None,
    }
    log(f"frame loop {res}")
    return res


def learned_vs_plain(args, cam, poses, views, L, h, odo, reset, plain_fps, plain_est, ref64) -> dict | None:
    """The learned form on the same sequence: per frame the same init calls,
    then the network input from the handle's level-2 buffers, the network
    (bf16 MFMA) and a seeded call with a short schedule and the keep-or-fall-
    back rule, all enqueued on the handle's stream.  Reports its frames/s next
    to the plain form's, its pose error against the FP64 oracle (plain init,
    full schedule) on the CPU sample, and -- on a fast-motion sequence with
    bursts up to 8 cm / 6 deg per frame -- both forms' error against the
    rendered ground truth."""
    import torch
    from slam_amd import synth, warmstart as ws
    if not os.path.exists(ws.WEIGHTS):
        log("learned: no trained weights; skipped")
        return None
    iters = [int(v) for v in args.learned_iters.split(",")]
    net = ws.WarmStartNet(max_batch=1)
    net.load()
    net.set_stream(odo.stream)
    x = torch.zeros((1, 60, 80, 4), dtype=torch.bfloat16, device="cuda")
    pred = torch.zeros((1, 6), dtype=torch.float32, device="cuda")
    it = np.ascontiguousarray(iters, dtype=np.int32)
    by = ctypes.byref

    def run_learned(k0, k1, vw):
        for k in range(k0, k1):
            L.odom_init_icp_model_current(h, by(vw["v4"][k - 1]), by(vw["n4"][k - 1]), CUTOFF)
            L.odom_init_rgb_model(h, by(vw["rgba"][k - 1]))
            L.odom_init_icp_raw_depth(h, by(vw["depth"][k]), MAXD, CUTOFF)
            L.odom_init_rgb(h, by(vw["rgb"][k]))
            L.odom_init_warmstart_input(h, x.data_ptr())
            ws._lib().cnn_forward(net.handle, x.data_ptr(), 1, pred.data_ptr())
            L.odom_track_seeded(h, pred.data_ptr(), ctypes.c_float(ws.T_SCALE), ctypes.c_float(ws.R_SCALE),
                                it.ctypes.data, 0, ctypes.c_float(ICP_WEIGHT), 1)

    nsteps = args.steps

    def pose12(p):
        T = np.eye(4)
        T[:3, :4] = p.reshape(3, 4)
        return T

    def timed(schedule) -> dict:
        it[:] = schedule
        reset()
        run_learned(1, 1 + args.warmup, views)
        odo.synchronize()
        log_t = torch.zeros((nsteps + 1, 12), dtype=torch.float32, device="cuda")
        odo.set_pose_log(log_t.data_ptr(), nsteps + 1)
        reset()
        torch.cuda.synchronize()
        odo.synchronize()
        s0 = odo.stats()
        t0 = time.perf_counter()
        run_learned(1, 1 + nsteps, views)
        odo.synchronize()
        el = time.perf_counter() - t0
        s1 = odo.stats()
        fps = nsteps / el
        est = log_t.cpu().numpy()[:nsteps]
        kept = (s1["seedKept"] - s0["seedKept"]) / max(1, s1["seedCalls"] - s0["seedCalls"])
        r = {"schedule": list(schedule), "frames_per_s": round(fps, 2), "speedup_vs_plain": round(fps / plain_fps, 3),
             "prediction_kept_frac": round(kept, 3)}
        if ref64 is not None:
            n = len(ref64)
            e = [synth.pose_error(pose12(est[k]), ref64[k]) for k in range(n)]
            r["vs_fp64_oracle_mean_mm"] = round(1e3 * float(np.mean([q[0] for q in e])), 4)
            r["vs_fp64_oracle_max_mm"] = round(1e3 * float(np.max([q[0] for q in e])), 4)
            r["vs_fp64_oracle_mean_deg"] = round(float(np.mean([q[1] for q in e])), 5)
        dt_l = [synth.pose_error(pose12(est[k - 1]), poses[k]) for k in range(1, nsteps + 1)]
        r["drift_vs_truth_mean_mm"] = round(1e3 * float(np.mean([d[0] for d in dt_l])), 3)
        return r

    out = {"workload": "same 640x480 sequence, network init + seeded ICP+RGB, no SO3, fallback to the previous "
                       "pose when the level-2 association from the prediction is worse",
           "short_schedule": timed(iters), "full_schedule": timed([10, 5, 4])}
    it[:] = iters

    # fast motion: bursts up to 8 cm / 6 deg per frame, both forms vs truth
    nb = args.burst_frames
    if nb > 1:
        bposes = synth.spline_trajectory(nb, seed=args.seed + 1, max_step_m=0.08, max_step_deg=6.0, bursts=True)
        scene = synth.Scene(args.seed + 1)
        bdev = {k: [] for k in ("depth", "rgb", "v4", "n4", "rgba")}
        for i, T in enumerate(bposes):
            f = scene.render(T, cam, noise_seed=5000 + i)
            bdev["depth"].append(torch.from_numpy(f.depth).cuda())
            bdev["rgb"].append(torch.from_numpy(f.rgb).cuda())
            bdev["v4"].append(torch.from_numpy(f.vert4).cuda())
            bdev["n4"].append(torch.from_numpy(f.norm4).cuda())
            bdev["rgba"].append(torch.from_numpy(f.rgba).cuda())
        from slam_amd.odometry import frame_view, FMT_DEPTH_U16, FMT_RGB8, FMT_RGBA8, FMT_RGBA32F
        bv = {k: [frame_view(t, fmt)[0] for t in bdev[k]] for k, fmt in
              (("depth", FMT_DEPTH_U16), ("rgb", FMT_RGB8), ("v4", FMT_RGBA32F), ("n4", FMT_RGBA32F),
               ("rgba", FMT_RGBA8))}
        res = {}
        for mode in ("plain", "learned"):
            T0 = bposes[0]
            odo.set_pose(T0[:3, 3], T0[:3, :3])
            odo.initFirstRGB(bdev["rgb"][0])
            lg = torch.zeros((nb, 12), dtype=torch.float32, device="cuda")
            odo.set_pose_log(lg.data_ptr(), nb)
            if mode == "plain":
                for k in range(1, nb):
                    L.odom_init_icp_model_current(h, by(bv["v4"][k - 1]), by(bv["n4"][k - 1]), CUTOFF)
                    L.odom_init_rgb_model(h, by(bv["rgba"][k - 1]))
                    L.odom_init_icp_raw_depth(h, by(bv["depth"][k]), MAXD, CUTOFF)
                    L.odom_init_rgb(h, by(bv["rgb"][k]))
                    L.odom_track_async(h, 0, ICP_WEIGHT, 1, 0, 1)
            else:
                run_learned(1, nb, bv)
            odo.synchronize()
            e = lg.cpu().numpy()[: nb - 1]
            errs = [synth.pose_error(pose12(e[k - 1]), bposes[k]) for k in range(1, nb)]
            # relative pose error per frame: estimated step vs true step
            rpe = [synth.pose_error(np.linalg.inv(pose12(e[k - 2])) @ pose12(e[k - 1]),
                                    np.linalg.inv(bposes[k - 1]) @ bposes[k])[0] for k in range(2, nb)]
            from slam_amd import evaluate as ev
            st = [k / 30.0 for k in range(1, nb)]
            aligned = ev.ate(ev.trajectory(st, bposes[1:nb]), ev.trajectory(st, [pose12(e[k - 1]) for k in range(1, nb)]))
            res[mode] = {"ate_mean_mm": round(1e3 * float(np.mean([q[0] for q in errs])), 2),
                         "ate_rmse_aligned_mm": round(1e3 * aligned["rmse"], 2),
                         "rpe_mean_mm": round(1e3 * float(np.mean(rpe)), 3),
                         "frames_lost_gt_5cm": int(sum(1 for r in rpe if r > 0.05))}
        steps = [synth.pose_error(bposes[k], bposes[k - 1]) for k in range(1, nb)]
        out["bursts"] = {"frames": nb, "max_step_mm": round(1e3 * max(q[0] for q in steps), 1),
                         "max_step_deg": round(max(q[1] for q in steps), 2), **{k: v for k, v in res.items()}}
    log(f"learned {out}")
    odo.set_pose_log(0, 0)
    net.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=499, help="frames tracked in the timed region")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-frames", type=int, default=20, help="frames of the CPU baseline / FP64 oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=50)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 64-pair measurement")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start measurement")
    ap.add_argument("--learned-iters", default="5,2,2", help="schedule of the learned form (levels 0,1,2)")
    ap.add_argument("--burst-frames", type=int, default=150)
    ap.add_argument("--no-graph", action="store_true", help="launch tracking calls kernel by kernel")
    ap.add_argument("--pose-log", default=None, help="write the timed frames' poses in the frame loop's log format")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 4 device copies of 256 MiB (the reference kernel of counter runs)")
    ap.add_argument("--batch-reps", type=int, default=20)
    ap.add_argument("--no-splat", action="store_true", help="skip the splat prediction measurement")
    ap.add_argument("--no-fusion", action="store_true", help="skip the surfel fusion measurement")
    ap.add_argument("--no-loop", action="store_true", help="skip the full frame-loop sequences")
    ap.add_argument("--loop-only", action="store_true", help="run only the full frame-loop sequences")
    ap.add_argument("--loop-frames", type=int, default=2000)
    ap.add_argument("--fusion-only", action="store_true", help="measure only the surfel fusion (profiling)")
    ap.add_argument("--fusion-frames", type=int, default=10)
    ap.add_argument("--fusion-reps", type=int, default=20)
    ap.add_argument("--splat-only", action="store_true", help="measure only the splat prediction (profiling)")
    ap.add_argument("--splat-surfels", type=int, default=2_000_000)
    ap.add_argument("--splat-reps", type=int, default=50)
    ap.add_argument("--batch-oracle-pairs", type=int, default=6,
                    help="pairs of the batch checked against the FP64 oracle on the host")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU: the odometry path runs one frame pair per "
                                   "GPU and shards by sequence; run with --gpus 1"}))
        return 2
    if args.calibrate_copy:
        # four 256 MiB device copies: the known bytes that scale the HBM
        # counters of a profiled run
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 4 x 256 MiB device copies")
        del src, dst
    if args.splat_only:
        print(json.dumps({"splat": splat_prediction(args)}), flush=True)
        return 0
    if args.fusion_only:
        print(json.dumps({"fusion": surfel_fusion(args)}), flush=True)
        return 0
    if args.loop_only:
        print(json.dumps({"frame_loop": frame_loop_sequence(args)}), flush=True)
        return 0

    import torch
    from slam_amd import synth
    from slam_amd.odometry import RGBDOdometry, device_info, lib, frame_view, FMT_DEPTH_U16, FMT_RGB8, \
        FMT_RGBA8, FMT_RGBA32F

    name, arch = device_info()
    log(f"device {name} {arch}")
    cam = synth.CAM_640
    nframes = max(args.steps, args.warmup) + 1
    scene = synth.Scene(args.seed)
    poses = synth.spline_trajectory(nframes, seed=args.seed)
    keep_host = args.cpu_frames + 1 if not args.no_cpu else 0
    dev = {k: [] for k in ("depth", "rgb", "v4", "n4", "rgba")}
    host = []
    t0 = time.time()
    for i, T in enumerate(poses):
        f = scene.render(T, cam, noise_seed=1000 + i)
        dev["depth"].append(torch.from_numpy(f.depth).cuda())
        dev["rgb"].append(torch.from_numpy(f.rgb).cuda())
        dev["v4"].append(torch.from_numpy(f.vert4).cuda())
        dev["n4"].append(torch.from_numpy(f.norm4).cuda())
        dev["rgba"].append(torch.from_numpy(f.rgba).cuda())
        if i < keep_host:
            host.append(f)
        if (i + 1) % 100 == 0:
            log(f"rendered {i + 1}/{nframes} frames")
    torch.cuda.synchronize()
    log(f"inputs resident: {nframes} frames in {time.time() - t0:.1f}s")

    # prebuilt frame views: the timed loop is only C calls
    views = {k: [frame_view(t, fmt)[0] for t in dev[k]] for k, fmt in
             (("depth", FMT_DEPTH_U16), ("rgb", FMT_RGB8), ("v4", FMT_RGBA32F), ("n4", FMT_RGBA32F),
              ("rgba", FMT_RGBA8))}
    L = lib()
    odo = RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    h = odo.handle
    odo.set_graph(not args.no_graph)
    by = ctypes.byref

    def reset():
        T0 = poses[0]
        odo.set_pose(T0[:3, 3], T0[:3, :3])
        odo.initFirstRGB(dev["rgb"][0])

    def run(k0: int, k1: int) -> None:
        for k in range(k0, k1):
            L.odom_init_icp_model_current(h, by(views["v4"][k - 1]), by(views["n4"][k - 1]), CUTOFF)
            L.odom_init_rgb_model(h, by(views["rgba"][k - 1]))
            L.odom_init_icp_raw_depth(h, by(views["depth"][k]), MAXD, CUTOFF)
            L.odom_init_rgb(h, by(views["rgb"][k]))
            L.odom_track_async(h, 0, ICP_WEIGHT, 1, 0, 1)

    # shapes: autotune on the first frame's inputs
    reset()
    run(1, 2)
    odo.synchronize()
    odo.autotune(False, ICP_WEIGHT)
    shapes = {f"{'ab'[k]}{l}": list(odo.get_step_shape(k, l)) for k in range(2) for l in range(3)}
    log(f"autotuned shapes {shapes}")

    reset()
    run(1, 1 + args.warmup)
    odo.synchronize()

    pose_log = torch.zeros((nframes, 12), dtype=torch.float32, device="cuda")
    odo.set_pose_log(pose_log.data_ptr(), nframes)
    reset()
    torch.cuda.synchronize()
    odo.synchronize()
    t0 = time.perf_counter()
    run(1, 1 + args.steps)
    enqueued = time.perf_counter() - t0
    odo.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps / elapsed
    log(f"timed {args.steps} frames in {elapsed * 1e3:.1f} ms: {fps:.1f} frames/s "
        f"(host enqueue {enqueued * 1e3:.1f} ms)")
    est = pose_log.cpu().numpy()[: args.steps]
    stats = odo.stats()

    def pose12(p):
        T = np.eye(4)
        T[:3, :4] = p.reshape(3, 4)
        return T

    drift = [synth.pose_error(pose12(est[k - 1]), poses[k]) for k in range(1, args.steps + 1)]
    drift_t = float(np.mean([d[0] for d in drift]))
    drift_r = float(np.mean([d[1] for d in drift]))
    # the reference's benchmark measures: ATE after Horn alignment and the
    # one-frame RPE, at 30 Hz stamps; optionally the pose log it writes
    from slam_amd import dataio
    from slam_amd import evaluate as ev
    stamps = [k / 30.0 for k in range(1, args.steps + 1)]
    est_poses = [pose12(est[k - 1]) for k in range(1, args.steps + 1)]
    gt_traj = ev.trajectory(stamps, [poses[k] for k in range(1, args.steps + 1)])
    est_traj = ev.trajectory(stamps, est_poses)
    ate_r = ev.ate(gt_traj, est_traj)
    rpe_r = ev.rpe(gt_traj, est_traj, delta=1, delta_unit="f")
    trajectory_eval = {"ate_rmse_mm": round(1e3 * ate_r["rmse"], 3), "ate_max_mm": round(1e3 * ate_r["max"], 3),
                       "rpe_rmse_mm": round(1e3 * rpe_r["translational"]["rmse"], 4),
                       "rpe_rmse_deg": round(rpe_r["rotational_deg"]["rmse"], 5),
                       "evaluator": "slam_amd/evaluate.py (benchmark/evaluate_ate.py, evaluate_rpe.py: "
                                    "Horn-aligned ATE, 1-frame RPE)"}
    if args.pose_log:
        dataio.write_pose_log(args.pose_log, list(range(1, args.steps + 1)), est_poses)

    # roofline of the step kernels: per level, the last iteration of a
    # completed call relaunched back to back between one HIP event pair on the
    # library's stream (odom_profile_steps); the call itself with an event
    # pair around every launch gives the end-to-end call time
    iters = [10, 5, 4]
    npx = [(cam.width >> l) * (cam.height >> l) for l in range(3)]
    odo.track_async(False, ICP_WEIGHT, True, False, True)
    odo.synchronize()
    steps = odo.profile_steps(False, ICP_WEIGHT, reps=args.profile_reps)
    us = {0: steps["us_a"], 1: steps["us_b"]}
    profs = [odo.profile_track(False, ICP_WEIGHT, True, False, True) for _ in range(3)]
    track_us = float(np.mean([p["us_total"] for p in profs]))
    tot_t = {k: sum(iters[l] * us[k][l] for l in range(3)) for k in (0, 1)}
    dom = 0 if tot_t[0] >= tot_t[1] else 1
    bpp = BYTES_A if dom == 0 else BYTES_B
    launches = sum(iters)
    bytes_per_launch = sum(iters[l] * bpp * npx[l] for l in range(3)) / launches
    avg_us = tot_t[dom] / launches
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    joint_bytes = sum(iters[l] * (BYTES_ICP + BYTES_RGB) * npx[l] for l in range(3))
    joint_gbs = joint_bytes / ((tot_t[0] + tot_t[1]) * 1e-6) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_step_a_traffic_640x480.json")
    if dom == 0 and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = int(json.load(fh)["traffic_bytes_per_launch_10_5_4"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_step_a (ICP products + photometric association)" if dom == 0 else
                "k_step_b (photometric products)",
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_us": round(avg_us, 2),
                "per_level_us": {"k_step_a": [round(x, 2) for x in us[0]], "k_step_b": [round(x, 2) for x in us[1]]},
                "joint_iteration_gbs": round(joint_gbs, 1), "track_call_us_events_per_launch": round(track_us, 1)}
    log(f"roofline {roofline}")

    cpu = None
    accuracy = None
    ref64 = None
    if not args.no_cpu and args.cpu_frames > 0:
        from oracle import oracle as orc
        n = min(args.cpu_frames, args.steps)

        def oracle_run(fp64: bool):
            o = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, fp64=fp64)
            T = poses[0].copy()
            o.initFirstRGB(host[0].rgb)
            out = []
            t_start = time.perf_counter()
            for k in range(1, n + 1):
                o.initICPModel(host[k - 1].vert4, host[k - 1].norm4, CUTOFF, T.astype(np.float32))
                o.initRGBModel(host[k - 1].rgba)
                o.initICP(orc.bilateral(host[k].depth, MAXD), CUTOFF)
                o.initRGB(host[k].rgb)
                t, R = o.getIncrementalTransformation(T[:3, 3], T[:3, :3], False, ICP_WEIGHT, True, False, True)
                T = synth.make_pose(R.astype(np.float64), t.astype(np.float64))
                out.append(T)
            return out, time.perf_counter() - t_start

        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ref32, t32 = oracle_run(False)
        log(f"cpu baseline (FP32 restatement, {cores} threads): {n} frames in {t32:.1f}s")
        ref64, t64 = oracle_run(True)
        errs = [synth.pose_error(pose12(est[k]), ref64[k]) for k in range(n)]
        errs32 = [synth.pose_error(pose12(est[k]), ref32[k]) for k in range(n)]
        accuracy = {"frames": n,
                    "vs_fp64_oracle_mean_mm": round(1e3 * float(np.mean([e[0] for e in errs])), 5),
                    "vs_fp64_oracle_max_mm": round(1e3 * float(np.max([e[0] for e in errs])), 5),
                    "vs_fp64_oracle_mean_deg": round(float(np.mean([e[1] for e in errs])), 6),
                    "vs_fp64_oracle_max_deg": round(float(np.max([e[1] for e in errs])), 6),
                    "vs_fp32_oracle_max_mm": round(1e3 * float(np.max([e[0] for e in errs32])), 5)}
        cpu = {"value": round(n / t32, 3), "unit": "frames/s", "cores": cores, "kind": "port",
               "sample": f"first {n} frames of the same 640x480 sequence, same calls, FP32 C restatement "
                         f"(oracle/odom_oracle.c) with OpenMP over rows"}
        log(f"accuracy {accuracy}")

    learned = None
    if not args.no_learned:
        learned = learned_vs_plain(args, cam, poses, views, L, h, odo, reset, fps, est, ref64)
    batched = None if args.no_batch else batched_pairs(args)
    splat_res = None if args.no_splat else splat_prediction(args)
    fusion_res = None if args.no_fusion else surfel_fusion(args)
    loop_res = None if args.no_loop else frame_loop_sequence(args)

    line = {
        "metric": "frames/s 640x480 ICP+RGB at <=1mm vs FP64 oracle; learned-init vs plain",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "640x480 ICP+RGB+SO3 frame-to-model tracking, 3 levels 10/5/4, icpWeight 10, "
                               "smooth spline <=1cm/0.5deg per frame, plain init (the previous pose, as the reference's frame loop)",
                   "frames": args.steps, "width": cam.width, "height": cam.height, "seed": args.seed,
                   "init": "plain"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "drift_vs_truth": {"mean_mm": round(1e3 * drift_t, 3), "mean_deg": round(drift_r, 4)},
        "trajectory": trajectory_eval,
        "last_stats": stats,
        "shapes": shapes,
        "learned": learned,
        "batched": batched,
        "splat": splat_res,
        "fusion": fusion_res,
        "frame_loop": loop_res,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
