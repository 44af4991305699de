#!/usr/bin/env python
"""Frames/s of dense RGB-D odometry on one MI355X.

Workload (the headline): a 640x480 frame-to-model sequence (seeded procedural
room, Perlin textures, depth noise 1 mm * z^2, 2 % dropout) on a smooth
spline of up to 1 cm / 0.5 deg per frame.  Each step tracks one frame exactly
as the reference's frame loop does (elastic_fusion_file.cpp:342-374):
bilateral filter, initICPModel (model maps of the previous frame at the
current estimate), initRGBModel, initICP, initRGB, then
getIncrementalTransformation with ICP+RGB (icpWeight 10), 3 levels at 10/5/4
iterations and the SO(3) pre-alignment.  Inputs are resident in HBM before
timing starts; every call is enqueued without host synchronisation and the
pose chain stays on the device.  Launch shapes come from the library's fixed
table, so the run does not depend on timing.

Prints one JSON line: throughput, the roofline of the dominant step kernel
(HIP events on the library's stream) against the 8 TB/s specification and a
copy peak measured in the same run, the CPU restatement's frames/s (1 core
and all cores, the 320x240 ICP-only and the 640x480 ICP+RGB configurations,
median and p90 after warm-up), the pose error against the FP64 oracle, and
the other configurations' legs: the learned start against the plain forms
under the same iteration policy, the 64-pair batch, the splat prediction,
the surfel fusion and the 8-sequence fast-motion frame loop.
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
# the iteration policy of the learned start and its plain control: a level
# ends once an update is below 0.1 mm and 0.1 mrad
CONV_T = 1e-4
CONV_R = 1e-4


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Status:
    """ORs the status codes of the C calls a timed loop makes; checked after
    the timing, so a refused call cannot silently shorten the work."""

    def __init__(self, what: str):
        self.what = what
        self.bad = 0
        self.first = None

    def __call__(self, rc: int) -> None:
        if rc != 0 and self.first is None:
            self.first = rc
        self.bad |= rc != 0

    def check(self) -> None:
        if self.bad:
            raise RuntimeError(f"{self.what}: a C call returned {self.first}")


def pose12(p):
    T = np.eye(4)
    T[:3, :4] = np.asarray(p).reshape(3, 4)
    return T


def stream_copy_peak(gib: float = 1.0, reps: int = 10) -> dict:
    """STREAM-style copy of `gib` GiB, device to device, timed with HIP
    events (bytes moved = 2 x size per copy): the library's copy kernel
    (odom_stream_copy: 16-byte accesses, four loads in flight per thread) and
    torch's device copy; the higher is the measured peak."""
    import torch
    from slam_amd.odometry import lib
    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ms = ctypes.c_float()
    rc = lib().odom_stream_copy(src.data_ptr(), dst.data_ptr(), n * 4, reps, ctypes.byref(ms), None)
    if rc != 0:
        raise RuntimeError(f"odom_stream_copy returned {rc}")
    assert torch.equal(src[:1024], dst[:1024]) and torch.equal(src[-1024:], dst[-1024:])
    kern = 2 * n * 4 / (ms.value * 1e-3) / 1e9
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    tms = e0.elapsed_time(e1) / reps
    tgbs = 2 * n * 4 / (tms * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return {"GBs": round(max(kern, tgbs), 1), "odom_stream_copy_GBs": round(kern, 1), "torch_copy_GBs": round(tgbs, 1),
            "bytes_per_copy": 2 * n * 4, "copies": reps}


class DeviceSequence:
    """A rendered sequence resident on the device with prebuilt frame views,
    so a timed loop is only C calls."""

    def __init__(self, cam, poses, scene, noise_base, keep_host=0, what="sequence"):
        import torch
        from slam_amd.odometry import frame_view, FMT_DEPTH_U16, FMT_RGB8, FMT_RGBA8, FMT_RGBA32F
        self.cam, self.poses = cam, poses
        self.dev = {k: [] for k in ("depth", "rgb", "v4", "n4", "rgba")}
        self.host = []
        t0 = time.time()
        for i, T in enumerate(poses):
            f = scene.render(T, cam, noise_seed=noise_base + i)
            self.dev["depth"].append(torch.from_numpy(f.depth).cuda())
            self.dev["rgb"].append(torch.from_numpy(f.rgb).cuda())
            self.dev["v4"].append(torch.from_numpy(f.vert4).cuda())
            self.dev["n4"].append(torch.from_numpy(f.norm4).cuda())
            self.dev["rgba"].append(torch.from_numpy(f.rgba).cuda())
            if i < keep_host:
                self.host.append(f)
        torch.cuda.synchronize()
        self.views = {k: [frame_view(t, fmt)[0] for t in self.dev[k]] for k, fmt in
                      (("depth", FMT_DEPTH_U16), ("rgb", FMT_RGB8), ("v4", FMT_RGBA32F), ("n4", FMT_RGBA32F),
                       ("rgba", FMT_RGBA8))}
        log(f"{what}: {len(poses)} frames {cam.width}x{cam.height} resident in {time.time() - t0:.1f}s")

    def feed(self, L, h, k, st: Status, icp_only=False) -> None:
        """The init calls of frame k against frame k-1's model maps at the
        handle's current estimate."""
        by = ctypes.byref
        v = self.views
        st(L.odom_init_icp_model_current(h, by(v["v4"][k - 1]), by(v["n4"][k - 1]), CUTOFF))
        if not icp_only:
            st(L.odom_init_rgb_model(h, by(v["rgba"][k - 1])))
        st(L.odom_init_icp_raw_depth(h, by(v["depth"][k]), MAXD, CUTOFF))
        if not icp_only:
            st(L.odom_init_rgb(h, by(v["rgb"][k])))


class Forms:
    """The tracking calls compared on one handle: the reference's, and the
    iteration policy with the constant-velocity start or the learned start
    (network input, forward pass and the judged call, all on the handle's
    stream)."""

    def __init__(self, L, odo, net=None):
        from slam_amd.odometry import JUDGE_COARSE, JUDGE_FULL, START_VELOCITY, Policy
        import torch
        self.L, self.odo, self.h = L, odo, odo.handle
        self.net = net
        self.pol_ref_conv = Policy.make((10, 5, 4), CONV_T, CONV_R, so3=True)
        self.pol_cv = Policy.make((10, 5, 4), CONV_T, CONV_R, start=START_VELOCITY, so3=False)
        self.pol_learned = Policy.make((10, 5, 4), CONV_T, CONV_R, judge=JUDGE_COARSE)
        self.pol_learned_full = Policy.make((10, 5, 4), CONV_T, CONV_R, judge=JUDGE_FULL)
        self.x = torch.zeros((1, 60, 80, 4), dtype=torch.bfloat16, device="cuda")
        self.pred = torch.zeros((1, 6), dtype=torch.float32, device="cuda")

    def call(self, form: str, st: Status) -> None:
        from slam_amd import warmstart as ws
        L, h, by = self.L, self.h, ctypes.byref
        if form == "reference":
            st(L.odom_track_async(h, 0, ICP_WEIGHT, 1, 0, 1))
        elif form == "reference_conv":
            st(L.odom_track_policy(h, None, 1.0, 1.0, by(self.pol_ref_conv), 0, ICP_WEIGHT))
        elif form == "plain_policy":
            st(L.odom_track_policy(h, None, 1.0, 1.0, by(self.pol_cv), 0, ICP_WEIGHT))
        elif form in ("learned", "learned_full"):
            st(L.odom_init_warmstart_input(h, self.x.data_ptr()))
            st(ws._lib().cnn_infer(self.net.handle, self.x.data_ptr(), 1, self.pred.data_ptr()))
            pol = self.pol_learned if form == "learned" else self.pol_learned_full
            st(L.odom_track_policy(h, self.pred.data_ptr(), ws.T_SCALE, ws.R_SCALE, by(pol), 0, ICP_WEIGHT))
        else:
            raise ValueError(form)


def run_form(seq: DeviceSequence, forms: Forms, form: str, nsteps: int, warmup: int, trace: bool = False) -> dict:
    """Frames 1..nsteps of the sequence with one form: warm-up frames, then
    the timed frames enqueued without host synchronisation (poses logged on
    the device).  With `trace`, the same frames again with the statistics
    read after every call (iterations, whether the prediction was kept) and
    the network's predictions; the runs are deterministic, so the trace
    repeats the timed run."""
    import torch
    L, odo, h = forms.L, forms.odo, forms.h
    T0 = seq.poses[0]

    def reset():
        odo.set_pose(T0[:3, 3], T0[:3, :3])
        odo.initFirstRGB(seq.dev["rgb"][0])

    st = Status(form)
    reset()
    for k in range(1, 1 + warmup):
        seq.feed(L, h, k, st)
        forms.call(form, st)
    odo.synchronize()
    log_t = torch.zeros((nsteps + 1, 12), dtype=torch.float32, device="cuda")
    odo.set_pose_log(log_t.data_ptr(), nsteps + 1)
    reset()
    torch.cuda.synchronize()
    odo.synchronize()
    t0 = time.perf_counter()
    for k in range(1, 1 + nsteps):
        seq.feed(L, h, k, st)
        forms.call(form, st)
    odo.synchronize()
    el = time.perf_counter() - t0
    st.check()
    est = [pose12(p) for p in log_t.cpu().numpy()[:nsteps]]
    odo.set_pose_log(0, 0)
    r = {"frames_per_s": round(nsteps / el, 2), "ms_per_frame": round(1e3 * el / nsteps, 4), "poses": est}
    if trace:
        reset()
        its, alts, kept, preds = [], [], [], []
        for k in range(1, 1 + nsteps):
            seq.feed(L, h, k, st)
            forms.call(form, st)
            s = odo.stats()
            its.append(s["gnIterations"])
            alts.append(s["altIterations"])
            kept.append(s["seedUsed"])
            if form.startswith("learned"):
                preds.append(forms.pred.cpu().numpy()[0].copy())
        st.check()
        r.update({"iters_per_frame": round(float(np.mean(its)), 3),
                  "alt_iters_per_frame": round(float(np.mean(alts)), 3), "_kept": kept, "_preds": preds})
        if form.startswith("learned"):
            r["fallback_rate"] = round(1.0 - float(np.mean(kept)), 4)
    return r


def error_stats(est, ref, key: str) -> dict:
    from slam_amd import synth
    e = [synth.pose_error(a, b) for a, b in zip(est, ref)]
    return {f"{key}_mean_mm": round(1e3 * float(np.mean([q[0] for q in e])), 5),
            f"{key}_max_mm": round(1e3 * float(np.max([q[0] for q in e])), 5),
            f"{key}_mean_deg": round(float(np.mean([q[1] for q in e])), 6),
            f"{key}_max_deg": round(float(np.max([q[1] for q in e])), 6)}


def oracle_sequence(seq: DeviceSequence, n: int, fp64: bool, call):
    """The oracle's tracker over frames 1..n of the sequence (host frames),
    with the per-frame call(o, T, k) -> (t, R, info)."""
    from oracle import oracle as orc
    from slam_amd import synth
    cam, host = seq.cam, seq.host
    o = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, fp64=fp64)
    T = seq.poses[0].copy()
    o.initFirstRGB(host[0].rgb)
    out, times = [], []
    for k in range(1, n + 1):
        t_start = time.perf_counter()
        o.initICPModel(host[k - 1].vert4, host[k - 1].norm4, CUTOFF, T.astype(np.float32))
        o.initRGBModel(host[k - 1].rgba)
        o.initICP(orc.bilateral(host[k].depth, MAXD), CUTOFF)
        o.initRGB(host[k].rgb)
        t, R, _ = call(o, T, k)
        times.append(time.perf_counter() - t_start)
        T = synth.make_pose(np.asarray(R, np.float64).reshape(3, 3), np.asarray(t, np.float64).reshape(3))
        out.append(T)
    return out, times


def reference_call(o, T, k):
    t, R = o.getIncrementalTransformation(T[:3, 3], T[:3, :3], False, ICP_WEIGHT, True, False, True)
    return t, R, None


def cpu_baseline(args, seq_c2: DeviceSequence) -> dict:
    """The reference algorithm's CPU path: the FP32 C restatement
    (oracle/odom_oracle.c, OpenMP over rows) on this host, in this process,
    on the same inputs -- 640x480 ICP+RGB+SO3 (the headline's first frames)
    and 320x240 ICP-only (a 1 cm / 0.5 deg sequence of the same kind of room)
    -- on 1 thread and on all threads: 3 warm-up frames, then the median and
    p90 of the per-frame time (bilateral filter, the init calls and
    getIncrementalTransformation)."""
    from oracle import oracle as orc
    from slam_amd import synth
    n = args.cpu_frames
    warm = 3
    hw = orc.hardware_concurrency()
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        nproc = os.cpu_count() or 1
    all_threads = orc.get_threads()
    omp_env = os.environ.get("OMP_NUM_THREADS")
    log(f"cpu baseline: nproc {nproc}, hardware_concurrency {hw}, OMP_NUM_THREADS {omp_env}, "
        f"all-core run on {all_threads} threads")

    # C1: 320x240 ICP-only, icpWeight 100 (RGBDOdometryef.cpp:276), no SO3
    cam1 = synth.CAM_320
    scene1 = synth.Scene(args.seed + 100)
    poses1 = synth.spline_trajectory(n + warm + 1, seed=args.seed + 100)
    host1 = [scene1.render(T, cam1, noise_seed=7000 + i) for i, T in enumerate(poses1)]

    def run_c1(threads):
        orc.set_threads(threads)
        o = orc.Odometry(cam1.width, cam1.height, cam1.cx, cam1.cy, cam1.fx, cam1.fy)
        T = poses1[0].copy()
        times = []
        for k in range(1, n + warm + 1):
            t_start = time.perf_counter()
            o.initICPModel(host1[k - 1].vert4, host1[k - 1].norm4, CUTOFF, T.astype(np.float32))
            o.initICP(orc.bilateral(host1[k].depth, MAXD), CUTOFF)
            t, R = o.getIncrementalTransformation(T[:3, 3], T[:3, :3], False, 100.0, True, False, False)
            times.append(time.perf_counter() - t_start)
            T = synth.make_pose(np.asarray(R, np.float64), np.asarray(t, np.float64))
        return times[warm:]

    def run_c2(threads):
        orc.set_threads(threads)
        _, times = oracle_sequence(seq_c2, n + warm, False, reference_call)
        return times[warm:]

    def summary(times, threads):
        med = float(np.median(times))
        return {"threads": threads, "frames": len(times), "median_ms": round(1e3 * med, 3),
                "p90_ms": round(1e3 * float(np.percentile(times, 90)), 3), "frames_per_s": round(1.0 / med, 3)}

    out = {"nproc": nproc, "hardware_concurrency": hw, "OMP_NUM_THREADS": omp_env, "warmup_frames": warm}
    for name, fn in (("c1_320x240_icp", run_c1), ("c2_640x480_icp_rgb_so3", run_c2)):
        out[name] = {}
        for label, threads in (("1_core", 1), ("all_cores", all_threads)):
            t0 = time.time()
            out[name][label] = summary(fn(threads), threads)
            log(f"cpu baseline {name} {label}: {out[name][label]} ({time.time() - t0:.1f}s)")
    orc.set_threads(all_threads)
    return out


def c1_gpu(args) -> dict:
    """The 320x240 ICP-only configuration on the GPU: the same sequence form
    as the headline (icpWeight 100, no SO3, 10/5/4), frames/s."""
    from slam_amd import synth
    from slam_amd.odometry import RGBDOdometry, lib
    cam = synth.CAM_320
    n = args.steps
    poses = synth.spline_trajectory(n + 1, seed=args.seed + 100)
    seq = DeviceSequence(cam, poses, synth.Scene(args.seed + 100), 7000, what="C1")
    L = lib()
    odo = RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    odo.set_graph(True)
    h = odo.handle
    st = Status("C1")

    def run():
        odo.set_pose(poses[0][:3, 3], poses[0][:3, :3])
        for k in range(1, n + 1):
            seq.feed(L, h, k, st, icp_only=True)
            st(L.odom_track_async(h, 0, 100.0, 1, 0, 0))
        odo.synchronize()

    run()
    t0 = time.perf_counter()
    run()
    el = time.perf_counter() - t0
    st.check()
    odo.close()
    return {"workload": "320x240 ICP-only frame-to-model, icpWeight 100, no SO3, 10/5/4, 1 cm / 0.5 deg spline",
            "frames_per_s": round(n / el, 2), "ms_per_frame": round(1e3 * el / n, 4), "frames": n}


def network_timing(reps: int = 50) -> dict:
    """Microseconds per call of the warm-start network: the fused one-launch
    inference (cnn_infer) and the training forward (cnn_forward, one GEMM
    launch per layer plus im2col / split-K passes), at batch 1 and 8, timed
    with HIP events over `reps` back-to-back calls."""
    import torch
    from slam_amd import warmstart as ws
    net = ws.WarmStartNet(max_batch=8)
    net.load()
    out = {}
    for n in (1, 8):
        x = torch.rand((n, 60, 80, 4), device="cuda").to(torch.bfloat16).contiguous()
        y = torch.empty((n, 6), dtype=torch.float32, device="cuda")
        for name, fn in (("infer", ws._lib().cnn_infer), ("forward", ws._lib().cnn_forward)):
            for _ in range(3):
                fn(net.handle, x.data_ptr(), n, y.data_ptr())
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                fn(net.handle, x.data_ptr(), n, y.data_ptr())
            e1.record()
            e1.synchronize()
            out[f"{name}_batch{n}"] = round(1e3 * e0.elapsed_time(e1) / reps, 2)
    net.close()
    return out


FORMS = ("reference", "reference_conv", "plain_policy", "learned", "learned_full")
FORM_DOC = {
    "reference": "previous-pose start, SO3 pre-alignment, fixed 10/5/4 (getIncrementalTransformation)",
    "reference_conv": "previous-pose start, SO3, 10/5/4 with the convergence test",
    "plain_policy": "constant-velocity start, no SO3, 10/5/4 with the convergence test (the learned form "
                    "without the network)",
    "learned": "network start (bf16 MFMA, batch 1, two launches) judged against the constant-velocity start "
               "after the level-2 refinement of both, convergence test",
    "learned_full": "as learned, both starts refined through every level before the judge",
}


def learned_vs_plain(args, seq_c2: DeviceSequence, L, odo, ref64) -> dict | None:
    """The learned start against the plain forms, on the headline's sequence
    and on a fast-motion sequence (bursts up to 8 cm / 6 deg per frame, the
    warm-start network's training motions): frames/s, Gauss-Newton
    iterations per frame (the kept refinement's, and the judged-away one's),
    the fall-back rate, and the pose error against the FP64 oracle of the
    reference schedule and of the same policy (the oracle fed the network's
    predictions the device used)."""
    from slam_amd import evaluate as ev, synth, warmstart as ws
    if not os.path.exists(ws.WEIGHTS):
        log("learned: no trained weights; skipped")
        return None
    net = ws.WarmStartNet(max_batch=1)
    net.load()
    net.set_stream(odo.stream)
    forms = Forms(L, odo, net)
    nsteps, warm = args.steps, args.warmup
    res = {"forms": FORM_DOC, "convergence_test": {"dt_m": CONV_T, "dw_rad": CONV_R}, "judge_tol": 0.01,
           "network_us": network_timing()}
    log(f"learned: network {res['network_us']}")

    def public(r):
        return {k: v for k, v in r.items() if k not in ("poses", "_kept", "_preds")}

    # ---- the headline's sequence (smooth, up to 1 cm / 0.5 deg per frame)
    c2 = {f: run_form(seq_c2, forms, f, nsteps, warm, trace=True) for f in FORMS}
    if ref64 is not None:
        m = min(len(ref64), nsteps)
        for f in FORMS:
            c2[f].update(error_stats(c2[f]["poses"][:m], ref64[:m], "vs_fp64_reference"))
            e = c2[f]
            c2[f]["accuracy_gate_1mm_0.1deg"] = bool(e["vs_fp64_reference_mean_mm"] <= 1.0 and
                                                      e["vs_fp64_reference_mean_deg"] <= 0.1)
        # the same policy in FP64 (the oracle's restatement, the device's predictions)
        mp = min(args.policy_oracle_frames, m)
        calls = {
            "reference_conv": lambda o, T, k: o.track_policy(T[:3, 3], T[:3, :3], conv=(CONV_T, CONV_R), so3=True),
            "plain_policy": lambda o, T, k: o.track_policy(T[:3, 3], T[:3, :3], conv=(CONV_T, CONV_R),
                                                           start="velocity"),
            "learned": lambda o, T, k: o.track_policy(T[:3, 3], T[:3, :3], c2["learned"]["_preds"][k - 1],
                                                      ws.T_SCALE, ws.R_SCALE, conv=(CONV_T, CONV_R), judge="coarse"),
            "learned_full": lambda o, T, k: o.track_policy(T[:3, 3], T[:3, :3], c2["learned_full"]["_preds"][k - 1],
                                                           ws.T_SCALE, ws.R_SCALE, conv=(CONV_T, CONV_R),
                                                           judge="full"),
        }
        for f, call in calls.items():
            t0 = time.time()
            ref_p, _ = oracle_sequence(seq_c2, mp, True, call)
            c2[f].update(error_stats(c2[f]["poses"][:mp], ref_p, "vs_fp64_same_policy"))
            log(f"learned: {f} vs the FP64 oracle of its policy on {mp} frames ({time.time() - t0:.1f}s)")
    base = c2["reference"]["frames_per_s"]
    ctrl = c2["plain_policy"]["frames_per_s"]
    for f in FORMS:
        c2[f]["speedup_vs_reference"] = round(c2[f]["frames_per_s"] / base, 3)
        c2[f]["speedup_vs_plain_policy"] = round(c2[f]["frames_per_s"] / ctrl, 3)
    res["smooth_640x480"] = {"workload": f"the headline's sequence, frames 1..{nsteps}",
                             **{f: public(c2[f]) for f in FORMS}}
    log(f"learned smooth {res['smooth_640x480']}")

    # ---- fast motion: bursts up to 8 cm / 6 deg per frame
    nb = args.burst_frames
    if nb > 1:
        cam = seq_c2.cam
        bposes = synth.spline_trajectory(nb, seed=args.seed + 1, max_step_m=0.08, max_step_deg=6.0, bursts=True)
        bseq = DeviceSequence(cam, bposes, synth.Scene(args.seed + 1), 5000, what="bursts")
        steps = [synth.pose_error(bposes[k], bposes[k - 1]) for k in range(1, nb)]
        out = {"workload": f"{nb} frames 640x480, bursts up to {1e3 * max(q[0] for q in steps):.1f} mm / "
                           f"{max(q[1] for q in steps):.2f} deg per frame",
               "frames": nb}
        st = [k / 30.0 for k in range(1, nb)]
        gt = ev.trajectory(st, bposes[1:nb])
        for f in FORMS:
            r = run_form(bseq, forms, f, nb - 1, min(warm, nb - 1), trace=True)
            est = r["poses"]
            errs = [synth.pose_error(est[k - 1], bposes[k]) for k in range(1, nb)]
            rpe = [synth.pose_error(np.linalg.inv(est[k - 2]) @ est[k - 1],
                                    np.linalg.inv(bposes[k - 1]) @ bposes[k])[0] for k in range(2, nb)]
            r.update({"ate_mean_mm": round(1e3 * float(np.mean([q[0] for q in errs])), 2),
                      "ate_rmse_aligned_mm": round(1e3 * ev.ate(gt, ev.trajectory(st, est))["rmse"], 2),
                      "rpe_mean_mm": round(1e3 * float(np.mean(rpe)), 3),
                      "frames_lost_gt_5cm": int(sum(1 for q in rpe if q > 0.05))})
            out[f] = public(r)
        base_b = out["reference"]["frames_per_s"]
        for f in FORMS:
            out[f]["speedup_vs_reference"] = round(out[f]["frames_per_s"] / base_b, 3)
        res["bursts_640x480"] = out
        log(f"learned bursts {out}")
        del bseq
    net.close()
    return res


BATCH_LAYOUT = (("CAM_320", 16), ("CAM_640", 32), ("CAM_848", 16))


def batched_pairs(args, peak_copy=None) -> dict:
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

    # launch shapes: the library's table row for batches
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
    l0_joint = (BYTES_ICP + BYTES_RGB) * npx[0] / ((us_a[0] + us_b[0]) * 1e-6) / 1e9
    traffic = joint_ratio = None
    tpath = os.path.join(ROOT, "profiles", "k_step_batch_traffic_64pairs.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        traffic = int(tj["a_traffic_bytes_per_launch_10_5_4"])
        joint_ratio = tj["joint_ratio_10_5_4"]
    out = {"workload": "64 pairs: 16x320x240 + 32x640x480 + 16x848x480, <=3cm/2deg, ICP+RGB+SO3 10/5/4, "
                       "one batched call per step (hipGraph)",
           "pairs_per_s": round(len(odos) / call_s, 1), "ms_per_call": round(call_s * 1e3, 3),
           "roofline": {"bound": "hbm", "kernel": "k_step_a (batched)", "achieved": round(ach, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "peak_measured": peak_copy, "frac_measured": round(ach / peak_copy, 4) if peak_copy else None,
                        "level0_GBs": round(l0_gbs, 1), "bytes_per_launch": int(bytes_a),
                        "avg_launch_us": round(avg_a, 2), "traffic": traffic,
                        "joint_traffic_vs_62B": joint_ratio},
           "per_level_us": {"k_step_a": [round(x, 2) for x in us_a], "k_step_b": [round(x, 2) for x in us_b]},
           "joint_iteration_gbs": round(joint, 1),
           "joint_frac": round(joint / HBM_PEAK_GBS, 4),
           "joint_frac_measured": round(joint / peak_copy, 4) if peak_copy else None,
           "level0_joint_gbs": round(l0_joint, 1),
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
    """The reference's frame loop without loop closure on the device, for n
    sequences in lockstep (slam_amd/frame_loop.py BatchFrameLoop): per frame
    each sequence's bilateral filter, its init calls, one batched tracking
    call for all of them (with the learned start: one network forward of
    batch n first), then each sequence's fusion, cleaning and next
    prediction, with the reference configs' values (depth cutoff 3 m,
    confidence 10, timeDelta 200).  Eight 640x480 sequences of seeded rooms
    with bursts up to 10 cm / 8 deg per frame, frames resident on the device;
    reports the whole loop's sequence-frames/s (host reads of each pose and
    surfel count included, as the reference loop makes them), the model
    sizes, ATE / RPE against the generated ground truth, the fall-back rate
    and the Gauss-Newton iterations per frame, plain against learned."""
    import torch
    from slam_amd import evaluate, synth
    from slam_amd.frame_loop import BatchFrameLoop, FrameLoop, LoopParams
    from slam_amd.odometry import JUDGE_COARSE

    cam = synth.CAM_640
    n, ns = args.loop_frames, args.loop_seqs
    seqs = []
    t0 = time.time()
    maxdt = maxdr = 0.0
    for i in range(ns):
        seed = 920 + i
        poses = synth.spline_trajectory(n, seed=seed, max_step_m=0.10, max_step_deg=8.0, bursts=True)
        steps = [synth.pose_error(poses[k + 1], poses[k]) for k in range(n - 1)]
        maxdt = max(maxdt, max(q[0] for q in steps))
        maxdr = max(maxdr, max(q[1] for q in steps))
        T0i = np.linalg.inv(poses[0])
        scene = synth.Scene(seed)
        frames = []
        for k, T in enumerate(poses):
            f = scene.render(T, cam, noise_seed=92000 + 5000 * i + k, model=False)
            frames.append((torch.from_numpy(f.rgb).cuda(), torch.from_numpy(f.depth).cuda()))
        seqs.append(([T0i @ T for T in poses], frames))
        log(f"frame loop: sequence {i + 1}/{ns} resident ({time.time() - t0:.1f}s)")
    torch.cuda.synchronize()

    def run(learned: bool) -> dict:
        p = LoopParams(learned=learned, conv_t=CONV_T if learned else 0.0, conv_r=CONV_R if learned else 0.0,
                       judge=JUDGE_COARSE, max_surfels=args.loop_surfels)
        loop = BatchFrameLoop(ns, cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, p)
        warm = min(5, n)
        for k in range(warm):
            loop.process([s[1][k][0] for s in seqs], [s[1][k][1] for s in seqs])
        torch.cuda.synchronize()
        t1 = time.time()
        for k in range(warm, n):
            loop.process([s[1][k][0] for s in seqs], [s[1][k][1] for s in seqs])
        torch.cuda.synchronize()
        el = time.time() - t1
        stamps = [float(k) for k in range(n)]
        per = []
        for i, (gt, _) in enumerate(seqs):
            gtt = evaluate.trajectory(stamps, gt)
            est = evaluate.trajectory(stamps, [T.astype(np.float64) for T in loop.poses[i]])
            ate = evaluate.ate(gtt, est)
            rpe = evaluate.rpe(gtt, est, delta=1, delta_unit="f")
            lost = [k for k, (g, e) in enumerate(zip(gt, loop.poses[i]))
                    if np.linalg.norm((np.linalg.inv(g) @ e.astype(np.float64))[:3, 3]) > 0.05]
            c = loop.counters(i)
            per.append({"ate_rmse_mm": round(1e3 * ate["rmse"], 2), "ate_max_mm": round(1e3 * ate["max"], 2),
                        "rpe_rmse_mm": round(1e3 * rpe["translational"]["rmse"], 3),
                        "rpe_rmse_deg": round(rpe["rotational_deg"]["rmse"], 4),
                        "frames_lost_gt_5cm": len(lost), "first_lost_frame": lost[0] if lost else None,
                        "surfels_final": loop.models[i].lastCount(),
                        "iters_per_frame": round(c["gn_iterations"] / max(1, c["frames_tracked"]), 3),
                        "alt_iters_per_frame": round(c["alt_iterations"] / max(1, c["frames_tracked"]), 3),
                        "fallback_rate": (round(1.0 - c["predictions_kept"] / c["predictions"], 4)
                                          if c["predictions"] else None)})
        loop.close()
        agg = {"sequence_frames_per_s": round(ns * (n - warm) / el, 1),
               "ms_per_frame_of_all_sequences": round(1e3 * el / (n - warm), 3),
               "ate_rmse_mm_median": float(np.median([q["ate_rmse_mm"] for q in per])),
               "ate_rmse_mm_mean": round(float(np.mean([q["ate_rmse_mm"] for q in per])), 2),
               "rpe_rmse_mm_median": float(np.median([q["rpe_rmse_mm"] for q in per])),
               "frames_lost_gt_5cm_total": int(sum(q["frames_lost_gt_5cm"] for q in per)),
               "iters_per_frame": round(float(np.mean([q["iters_per_frame"] for q in per])), 3),
               "alt_iters_per_frame": round(float(np.mean([q["alt_iters_per_frame"] for q in per])), 3)}
        if learned:
            agg["fallback_rate"] = round(float(np.mean([q["fallback_rate"] for q in per])), 4)
        log(f"frame loop {'learned' if learned else 'plain'}: {agg}")
        return {**agg, "per_sequence": per}

    res = {"workload": f"{ns} sequences x {n} frames 640x480 (frame loop: track, fuse, clean, predict), bursts up to "
                       f"{1e3 * maxdt:.0f} mm / {maxdr:.1f} deg per frame, no loop closure; one batched tracking "
                       f"call per frame",
           "sequences": ns, "frames": n, "max_step_mm": round(1e3 * maxdt, 1), "max_step_deg": round(maxdr, 2),
           "plain": run(False)}
    from slam_amd import warmstart as ws
    res["learned"] = run(True) if os.path.exists(ws.WEIGHTS) else None
    del seqs
    torch.cuda.empty_cache()

    # the loop's own speed on a sequence it tracks throughout (smooth, 300 frames)
    scene = synth.Scene(921)
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
    res["smooth"] = {"frames": len(sposes), "frames_per_s": round((len(sposes) - 5) / sel, 1),
                     "ms_per_frame": round(1e3 * sel / (len(sposes) - 5), 3),
                     "surfels_final": sloop.model.lastCount(),
                     "ate_rmse_mm": round(1e3 * evaluate.ate(sgt, sest)["rmse"], 2),
                     "workload": "one sequence, 300 frames on a <=2 cm / 1 deg spline, native loop"}
    sloop.close()
    del sframes
    log(f"frame loop {res['smooth']}")
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=499, help="frames tracked in the timed region")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-frames", type=int, default=20, help="frames of the CPU baseline / FP64 oracle sample")
    ap.add_argument("--policy-oracle-frames", type=int, default=8,
                    help="frames of each policy checked against the FP64 oracle of the same policy")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=50)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 64-pair measurement")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched 64-pair measurement")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start measurement")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 320x240 ICP-only measurement (whose level 0 has the grid of a 640x480 level 1)")
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
    ap.add_argument("--loop-seqs", type=int, default=8)
    ap.add_argument("--loop-surfels", type=int, default=2_500_000)
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
    if args.batch_only:
        print(json.dumps({"batched": batched_pairs(args)}), flush=True)
        return 0

    import torch
    from slam_amd import synth
    from slam_amd.odometry import RGBDOdometry, device_info, lib

    name, arch = device_info()
    log(f"device {name} {arch}")
    copy_peak = stream_copy_peak()
    log(f"copy peak {copy_peak}")
    peak_copy = copy_peak["GBs"]
    cam = synth.CAM_640
    nframes = max(args.steps, args.warmup) + 1
    poses = synth.spline_trajectory(max(nframes, args.cpu_frames + 5), seed=args.seed)
    keep_host = args.cpu_frames + 5 if not args.no_cpu else 0
    seq = DeviceSequence(cam, poses, synth.Scene(args.seed), 1000, keep_host=keep_host, what="headline")

    L = lib()
    odo = RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    odo.set_graph(not args.no_graph)
    shapes = {f"{'ab'[k]}{l}": list(odo.get_step_shape(k, l)) for k in range(2) for l in range(3)}
    log(f"launch shapes (library table) {shapes}")
    forms = Forms(L, odo)

    head = run_form(seq, forms, "reference", args.steps, args.warmup)
    fps = head["frames_per_s"]
    elapsed = args.steps / fps
    est = head["poses"]
    log(f"timed {args.steps} frames: {fps:.1f} frames/s")
    stats = odo.stats()

    drift = [synth.pose_error(est[k - 1], poses[k]) for k in range(1, args.steps + 1)]
    drift_t = float(np.mean([d[0] for d in drift]))
    drift_r = float(np.mean([d[1] for d in drift]))
    # the reference's benchmark measures: ATE after Horn alignment and the
    # one-frame RPE, at 30 Hz stamps; optionally the pose log it writes
    from slam_amd import dataio
    from slam_amd import evaluate as ev
    stamps = [k / 30.0 for k in range(1, args.steps + 1)]
    gt_traj = ev.trajectory(stamps, [poses[k] for k in range(1, args.steps + 1)])
    est_traj = ev.trajectory(stamps, est)
    ate_r = ev.ate(gt_traj, est_traj)
    rpe_r = ev.rpe(gt_traj, est_traj, delta=1, delta_unit="f")
    trajectory_eval = {"ate_rmse_mm": round(1e3 * ate_r["rmse"], 3), "ate_max_mm": round(1e3 * ate_r["max"], 3),
                       "rpe_rmse_mm": round(1e3 * rpe_r["translational"]["rmse"], 4),
                       "rpe_rmse_deg": round(rpe_r["rotational_deg"]["rmse"], 5),
                       "evaluator": "slam_amd/evaluate.py (benchmark/evaluate_ate.py, evaluate_rpe.py: "
                                    "Horn-aligned ATE, 1-frame RPE)"}
    if args.pose_log:
        dataio.write_pose_log(args.pose_log, list(range(1, args.steps + 1)), est)

    # roofline of the step kernels: per level, the last iteration of a
    # completed call relaunched back to back, one graph replay between one HIP
    # event pair on the library's stream (odom_profile_steps); the call itself with an event
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
                "peak_measured": peak_copy, "frac_measured": round(achieved / peak_copy, 4),
                "kernel": "k_step_a (ICP products + photometric association)" if dom == 0 else
                "k_step_b (photometric products)",
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_us": round(avg_us, 2),
                "per_level_us": {"k_step_a": [round(x, 2) for x in us[0]], "k_step_b": [round(x, 2) for x in us[1]]},
                "joint_iteration_gbs": round(joint_gbs, 1), "track_call_us_events_per_launch": round(track_us, 1)}
    log(f"roofline {roofline}")

    cpu = None
    accuracy = None
    ref64 = None
    cpu_detail = None
    if not args.no_cpu and args.cpu_frames > 0:
        n = min(args.cpu_frames, args.steps)
        cpu_detail = cpu_baseline(args, seq)
        c2 = cpu_detail["c2_640x480_icp_rgb_so3"]["all_cores"]
        cpu = {"value": c2["frames_per_s"], "unit": "frames/s", "cores": c2["threads"], "kind": "port",
               "sample": f"640x480 ICP+RGB+SO3 frames of the same sequence (3 warm-up, then the median of "
                         f"{c2['frames']}), the FP32 C restatement (oracle/odom_oracle.c) with OpenMP over rows "
                         f"on {c2['threads']} threads; 1 core: "
                         f"{cpu_detail['c2_640x480_icp_rgb_so3']['1_core']['frames_per_s']} frames/s",
               "median_ms": c2["median_ms"], "p90_ms": c2["p90_ms"], "detail": cpu_detail}
        ref64, t64 = oracle_sequence(seq, n, True, reference_call)
        ref32, _ = oracle_sequence(seq, n, False, reference_call)
        accuracy = {"frames": n, **error_stats(est[:n], ref64, "vs_fp64_oracle"),
                    "vs_fp32_oracle_max_mm": error_stats(est[:n], ref32, "x")["x_max_mm"]}
        log(f"accuracy {accuracy}")

    c1 = None if args.no_c1 else c1_gpu(args)
    log(f"C1 {c1}")
    learned = None if args.no_learned else learned_vs_plain(args, seq, L, odo, ref64)
    del seq
    torch.cuda.empty_cache()
    batched = None if args.no_batch else batched_pairs(args, peak_copy)
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
        "copy_peak": copy_peak,
        "drift_vs_truth": {"mean_mm": round(1e3 * drift_t, 3), "mean_deg": round(drift_r, 4)},
        "trajectory": trajectory_eval,
        "last_stats": stats,
        "shapes": shapes,
        "c1_gpu": c1,
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
