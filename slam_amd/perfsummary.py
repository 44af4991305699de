"""Summaries of rocprofv3 output for the committed profiles.

    python -m slam_amd.perfsummary stats  DIR [--top N] [--cmd "..."]
        markdown table of the kernel statistics (DIR/run_kernel_stats.csv)
    python -m slam_amd.perfsummary frame  DIR
        kernels of one mid-run tracking call (DIR/run_kernel_trace.csv)
    python -m slam_amd.perfsummary loop   DIR [--frames A B]
        frames A..B of the longest frame-loop run in the trace (a frame runs
        from one bilateral launch to the next and holds a tracking call and
        a splat prediction): kernel time per frame, the frame's span, and
        the idle gaps above 4 us with the launches on either side
    python -m slam_amd.perfsummary traffic FETCH_DIR WRITE_DIR [--out JSON]
        HBM bytes per launch of the ICP/association pass (k_step_a) per
        pyramid level, from FETCH_SIZE and WRITE_SIZE counter runs (KiB per
        dispatch), each scaled by what the counter reads for the 256 MiB
        device copies of ``bench.py --calibrate-copy``.
    python -m slam_amd.perfsummary batch-traffic FETCH_DIR WRITE_DIR [--out JSON]
        the same for the 64-pair batch's passes A (unfused) and B per level,
        with the joint (A + B) traffic against the 62 B/px contract
    python -m slam_amd.perfsummary kernel-traffic FETCH_DIR WRITE_DIR --kernel NAME --algorithmic BYTES
        the same for every dispatch of the kernels whose name contains NAME:
        mean FETCH_SIZE and WRITE_SIZE, copy-scaled bytes, and their ratio to
        BYTES (the algorithmic bytes per dispatch)
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
import sys
from collections import defaultdict

COPY_BYTES = 256 << 20
LEVEL_PX = (640 * 480, 320 * 240, 160 * 120)
MAX_BLK = 256


def short(name: str) -> str:
    n = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").replace("slam::", "").strip()
    return n


def _stats_file(d: str) -> str:
    """The kernel statistics of an output directory (``-o`` prefix any)."""
    names = sorted(f for f in os.listdir(d) if f.endswith("_kernel_stats.csv"))
    if not names:
        raise FileNotFoundError(f"no *_kernel_stats.csv in {d}")
    return os.path.join(d, "run_kernel_stats.csv" if "run_kernel_stats.csv" in names else names[0])


def cmd_stats(d: str, top: int, cmd: str | None) -> str:
    rows = list(csv.DictReader(open(_stats_file(d))))
    out = []
    if cmd:
        out.append(f"Command: `{cmd}`\n")
    out.append("| kernel | calls | total ms | avg us | min us | max us | % |")
    out.append("|---|---:|---:|---:|---:|---:|---:|")
    for r in rows[:top]:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                   f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def cmd_frame(d: str, mode: str = "plain") -> str:
    tr = list(csv.DictReader(open(os.path.join(d, "run_kernel_trace.csv"))))
    tr.sort(key=lambda r: int(r["Start_Timestamp"]))
    ends = [i for i, r in enumerate(tr) if "k_track_end" in r["Kernel_Name"]]
    frames = []
    for a, b in zip(ends[:-1], ends[1:]):
        names = " ".join(r["Kernel_Name"] for r in tr[a + 1:b + 1])
        if "k_batch" in names or b - a > 400:
            continue
        seeded = "k_seed" in names
        so3 = "k_so3_iter" in names or "k_so3_loop" in names
        batched = "k_solve" in names and not seeded     # the batch solves in launches of its own
        if mode == "any" or (mode == "plain" and so3 and not seeded and not batched) or (mode == "learned" and seeded):
            frames.append((a, b))
    i0, i1 = frames[len(frames) // 2]
    t0 = int(tr[i0]["End_Timestamp"])
    agg = defaultdict(lambda: [0, 0.0])
    for r in tr[i0 + 1:i1 + 1]:
        a = agg[short(r["Kernel_Name"])]
        a[0] += 1
        a[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    total = (int(tr[i1]["End_Timestamp"]) - t0) / 1e3
    out = ["| kernel | launches | us per frame | avg us |", "|---|---:|---:|---:|"]
    for k, (n, us) in sorted(agg.items(), key=lambda x: -x[1][1]):
        out.append(f"| `{k}` | {n} | {us:.1f} | {us / n:.2f} |")
    out.append(f"\nOne tracking call (frame) end to end: {total:.1f} us, {i1 - i0} launches.")
    return "\n".join(out)


def cmd_runs(d: str, kernel: str, min_len: int = 40) -> str:
    """Runs of at least min_len back-to-back launches of one step kernel (the
    relaunches of odom_profile_steps): their mean kernel duration, the figure
    the benchmark's event pair measures, and the mean spacing of their starts."""
    tr = list(csv.DictReader(open(os.path.join(d, "run_kernel_trace.csv"))))
    tr.sort(key=lambda r: int(r["Start_Timestamp"]))
    out = ["| kernel | launches | avg us | start-to-start us |", "|---|---:|---:|---:|"]
    i = 0
    while i < len(tr):
        j = i
        while j < len(tr) and tr[j]["Kernel_Name"] == tr[i]["Kernel_Name"]:
            j += 1
        if j - i >= min_len and kernel in tr[i]["Kernel_Name"]:
            dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in tr[i:j]]
            gap = (int(tr[j - 1]["Start_Timestamp"]) - int(tr[i]["Start_Timestamp"])) / 1e3 / (j - i - 1)
            out.append(f"| `{short(tr[i]['Kernel_Name'])}` | {j - i} | {sum(dur) / len(dur):.2f} | {gap:.2f} |")
        i = j
    return "\n".join(out)


def cmd_loop(d: str, first: int, last: int) -> str:
    tr = list(csv.DictReader(open(os.path.join(d, "run_kernel_trace.csv"))))
    tr.sort(key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(tr) if "k_bilateral" in r["Kernel_Name"]]
    # the longest run of consecutive single-sequence loop frames (each one
    # bilateral filter, tracking call and prediction); the batched loops
    # interleave their sequences' launches and do not form such runs
    runs, cur = [], []
    for a, b in zip(starts[:-1], starts[1:]):
        names = " ".join(r["Kernel_Name"] for r in tr[a:b])
        if "k_track_end" in names and "k_splat_raster" in names:
            cur.append((a, b))
        elif cur:
            runs.append(cur)
            cur = []
    if cur:
        runs.append(cur)
    frames = max(runs, key=len)[first:last] if runs else []
    if not frames:
        raise ValueError("no frame-loop frames in the trace")
    agg = defaultdict(lambda: [0, 0.0])
    gaps = defaultdict(lambda: [0, 0.0])
    span = busy = 0.0
    for a, b in frames:
        span += (int(tr[b]["Start_Timestamp"]) - int(tr[a]["Start_Timestamp"])) / 1e3
        for i in range(a, b):
            r = tr[i]
            us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            busy += us
            g = agg[short(r["Kernel_Name"])]
            g[0] += 1
            g[1] += us
            idle = (int(tr[i + 1]["Start_Timestamp"]) - int(r["End_Timestamp"])) / 1e3
            if idle > 4.0:
                e = gaps[f"{short(r['Kernel_Name'])} -> {short(tr[i + 1]['Kernel_Name'])}"]
                e[0] += 1
                e[1] += idle
    n = len(frames)
    out = [f"Frames {first}..{first + n - 1} of the longest run of single-sequence loop frames: span {span / n:.1f} us per frame, "
           f"kernels {busy / n:.1f} us per frame.\n",
           "| kernel | launches per frame | us per frame |", "|---|---:|---:|"]
    for k, (c, us) in sorted(agg.items(), key=lambda x: -x[1][1]):
        out.append(f"| `{k}` | {c / n:.2f} | {us / n:.1f} |")
    out += ["\nIdle gaps above 4 us:\n", "| after -> before | per frame | us per frame |", "|---|---:|---:|"]
    for k, (c, us) in sorted(gaps.items(), key=lambda x: -x[1][1]):
        out.append(f"| `{k}` | {c / n:.2f} | {us / n:.1f} |")
    return "\n".join(out)


def _level(name: str, grid: int, wg: int) -> int | None:
    m = re.search(r"k_step_a<(\d+), (\d+)", name)
    if not m:
        return None
    threads, ppt = int(m.group(1)), int(m.group(2))
    blocks = grid // wg
    for lv, n in enumerate(LEVEL_PX):
        if min(MAX_BLK, -(-n // (threads * ppt))) == blocks:
            return lv
    return None


def _calibration_copies(rows) -> list:
    """Counter values of the four 256 MiB calibration copies of
    ``bench.py --calibrate-copy``: the first four device copies of the run in
    dispatch order (the copy-peak measurement that may follow uses the same
    blit grid on 1 GiB buffers)."""
    cp = [r for r in rows if "copyBuffer" in r["Kernel_Name"] and int(r["Grid_Size"]) == 131072]
    if cp and "Dispatch_Id" in cp[0]:
        cp.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [float(r["Counter_Value"]) for r in cp[:4]]


def _per_level(path: str):
    rows = list(csv.DictReader(open(path)))
    copy = _calibration_copies(rows)
    lv = defaultdict(list)
    for r in rows:
        if "k_step_a" not in r["Kernel_Name"] or "true>" not in r["Kernel_Name"]:
            continue  # the fused pass of tracking calls only
        level = _level(r["Kernel_Name"], int(r["Grid_Size"]), int(r["Workgroup_Size"]))
        if level is not None:
            lv[level].append(float(r["Counter_Value"]))
    return copy, {k: sum(v) / len(v) for k, v in lv.items()}, {k: len(v) for k, v in lv.items()}


def cmd_traffic(fetch_dir: str, write_dir: str) -> dict:
    fc, fl, fn = _per_level(os.path.join(fetch_dir, "run_counter_collection.csv"))
    wc, wl, wn = _per_level(os.path.join(write_dir, "run_counter_collection.csv"))
    f_scale = COPY_BYTES / (1024.0 * (sum(fc) / len(fc)))
    w_scale = COPY_BYTES / (1024.0 * (sum(wc) / len(wc)))
    levels = {}
    for l in sorted(set(fl) & set(wl)):
        fb = fl[l] * 1024.0 * f_scale
        wb = wl[l] * 1024.0 * w_scale
        levels[str(l)] = {"fetch_kib": round(fl[l], 1), "write_kib": round(wl[l], 1), "dispatches": fn[l],
                          "fetch_bytes": int(fb), "write_bytes": int(wb), "traffic_bytes": int(fb + wb),
                          "algorithmic_bytes_48B": 48 * LEVEL_PX[l]}
    iters = (10, 5, 4)
    per_launch = sum(iters[int(l)] * v["traffic_bytes"] for l, v in levels.items()) / sum(
        iters[int(l)] for l in levels)
    return {"kernel": "k_step_a (fused pass of tracking calls)", "counter_unit": "KiB per dispatch",
            "calibration": {"copy_bytes": COPY_BYTES, "fetch_kib": round(sum(fc) / len(fc), 1),
                            "write_kib": round(sum(wc) / len(wc), 1), "fetch_scale": round(f_scale, 4),
                            "write_scale": round(w_scale, 4)},
            "levels": levels, "traffic_bytes_per_launch_10_5_4": int(per_launch)}


def _kernel_counter(path: str, name: str):
    rows = list(csv.DictReader(open(path)))
    copy = _calibration_copies(rows)
    vals = [float(r["Counter_Value"]) for r in rows if name in r["Kernel_Name"]]
    if not copy or not vals:
        raise ValueError(f"no copy calibration or no {name} dispatches in {path}")
    return sum(copy) / len(copy), sum(vals) / len(vals), len(vals)


def cmd_kernel_traffic(fetch_dir: str, write_dir: str, name: str, algorithmic: float) -> dict:
    fc, fv, fn = _kernel_counter(os.path.join(fetch_dir, "run_counter_collection.csv"), name)
    wc, wv, wn = _kernel_counter(os.path.join(write_dir, "run_counter_collection.csv"), name)
    fb = fv * COPY_BYTES / fc
    wb = wv * COPY_BYTES / wc
    return {"kernel": name, "counter_unit": "KiB per dispatch", "dispatches": [fn, wn],
            "calibration": {"copy_bytes": COPY_BYTES, "fetch_kib": round(fc, 1), "write_kib": round(wc, 1),
                            "fetch_scale": round(COPY_BYTES / (1024.0 * fc), 4),
                            "write_scale": round(COPY_BYTES / (1024.0 * wc), 4)},
            "fetch_kib": round(fv, 1), "write_kib": round(wv, 1), "fetch_bytes": int(fb), "write_bytes": int(wb),
            "traffic_bytes": int(fb + wb), "algorithmic_bytes": int(algorithmic),
            "ratio": round((fb + wb) / algorithmic, 3)}


# The 64-pair batch (16 at 320x240, 32 at 640x480, 16 at 848x480): pixels per
# level over all pairs, and the schedule's order of levels within a call.
BATCH_LEVEL_PX = (17_571_840, 4_392_960, 1_098_240)
CALL_LEVELS = [2] * 4 + [1] * 5 + [0] * 10


def _batch_levels(path: str, kernel: str):
    """Counter values of the batched calls' dispatches of `kernel` (the
    unfused pass A or pass B), by level: every batched call launches each
    pass 19 times in the schedule's order (4 at level 2, 5 at 1, 10 at 0)."""
    rows = list(csv.DictReader(open(path)))
    copy = _calibration_copies(rows)
    if "Dispatch_Id" in rows[0]:
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    if kernel == "a":
        sel = [r for r in rows if "k_step_a" in r["Kernel_Name"] and "false>" in r["Kernel_Name"]]
    else:
        sel = [r for r in rows if "k_step_b" in r["Kernel_Name"]]
    if not copy or not sel or len(sel) % len(CALL_LEVELS):
        raise ValueError(f"{path}: {len(copy)} calibration copies, {len(sel)} dispatches of pass {kernel}")
    lv = defaultdict(list)
    for i, r in enumerate(sel):
        lv[CALL_LEVELS[i % len(CALL_LEVELS)]].append(float(r["Counter_Value"]))
    return sum(copy) / len(copy), {k: sum(v) / len(v) for k, v in lv.items()}, {k: len(v) for k, v in lv.items()}


def cmd_batch_traffic(fetch_dir: str, write_dir: str) -> dict:
    """HBM bytes per dispatch of the batched passes A and B per level, copy-scaled,
    against the algorithmic bytes (A 58 B/px: the ICP maps and the photometric
    association reads; B 4 B/px: the gradients; joint 62 B/px)."""
    out = {"workload": "64 pairs (16 x 320x240, 32 x 640x480, 16 x 848x480), batched ICP+RGB+SO3 calls",
           "counter_unit": "KiB per dispatch", "levels": {}}
    per = {}
    for k, bpp in (("a", 58), ("b", 4)):
        fc, fl, fn = _batch_levels(os.path.join(fetch_dir, "run_counter_collection.csv"), k)
        wc, wl, wn = _batch_levels(os.path.join(write_dir, "run_counter_collection.csv"), k)
        fs, ws = COPY_BYTES / (1024.0 * fc), COPY_BYTES / (1024.0 * wc)
        out[f"calibration_{k}"] = {"fetch_kib": round(fc, 1), "write_kib": round(wc, 1),
                                   "fetch_scale": round(fs, 4), "write_scale": round(ws, 4)}
        for l in sorted(fl):
            fb, wb = fl[l] * 1024.0 * fs, wl[l] * 1024.0 * ws
            alg = bpp * BATCH_LEVEL_PX[l]
            per[(k, l)] = fb + wb
            out["levels"].setdefault(str(l), {})[k] = {
                "dispatches": fn[l], "fetch_kib": round(fl[l], 1), "write_kib": round(wl[l], 1),
                "fetch_bytes": int(fb), "write_bytes": int(wb), "traffic_bytes": int(fb + wb),
                "algorithmic_bytes": alg, "ratio": round((fb + wb) / alg, 3)}
    for l in range(3):
        if ("a", l) in per and ("b", l) in per:
            alg = 62 * BATCH_LEVEL_PX[l]
            out["levels"][str(l)]["joint"] = {"traffic_bytes": int(per[("a", l)] + per[("b", l)]),
                                              "algorithmic_bytes": alg,
                                              "ratio": round((per[("a", l)] + per[("b", l)]) / alg, 3)}
    iters = (10, 5, 4)
    out["a_traffic_bytes_per_launch_10_5_4"] = int(sum(iters[l] * per[("a", l)] for l in range(3)) / sum(iters))
    out["joint_ratio_10_5_4"] = round(sum(iters[l] * (per[("a", l)] + per[("b", l)]) for l in range(3)) /
                                      sum(iters[l] * 62 * BATCH_LEVEL_PX[l] for l in range(3)), 3)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=20)
    s.add_argument("--command", default=None, help="the command that made the profile")
    f = sub.add_parser("frame")
    f.add_argument("dir")
    f.add_argument("--mode", choices=("plain", "learned", "any"), default="plain",
                   help="plain: calls with the SO3 pass; learned: seeded calls")
    rn = sub.add_parser("runs")
    rn.add_argument("dir")
    rn.add_argument("--kernel", default="k_step_")
    rn.add_argument("--min", type=int, default=40)
    lp = sub.add_parser("loop")
    lp.add_argument("dir")
    lp.add_argument("--frames", type=int, nargs=2, default=(10, 200))
    kt = sub.add_parser("kernel-traffic")
    kt.add_argument("fetch_dir")
    kt.add_argument("write_dir")
    kt.add_argument("--kernel", required=True)
    kt.add_argument("--algorithmic", type=float, required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--out", default=None)
    bt = sub.add_parser("batch-traffic")
    bt.add_argument("fetch_dir")
    bt.add_argument("write_dir")
    bt.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.cmd == "stats":
        print(cmd_stats(a.dir, a.top, a.command))
    elif a.cmd == "frame":
        print(cmd_frame(a.dir, a.mode))
    elif a.cmd == "runs":
        print(cmd_runs(a.dir, a.kernel, a.min))
    elif a.cmd == "loop":
        print(cmd_loop(a.dir, *a.frames))
    elif a.cmd == "kernel-traffic":
        print(json.dumps(cmd_kernel_traffic(a.fetch_dir, a.write_dir, a.kernel, a.algorithmic), indent=1))
    else:
        res = cmd_traffic(a.fetch_dir, a.write_dir) if a.cmd == "traffic" else \
            cmd_batch_traffic(a.fetch_dir, a.write_dir)
        text = json.dumps(res, indent=1)
        if a.out:
            with open(a.out, "w") as fh:
                fh.write(text + "\n")
        print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
