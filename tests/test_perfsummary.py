"""The rocprofv3 summaries behind the committed profiles, on small synthetic
CSV files whose figures are worked out by hand: the copy calibration of the
counters (only the first four 256 MiB copies, by dispatch order), the
per-level assignment of batched dispatches, the relaunch runs and the choice
of a plain or learned frame from a kernel trace."""
import csv

import pytest

from slam_amd import perfsummary as ps

COUNTER_COLS = ["Correlation_Id", "Dispatch_Id", "Grid_Size", "Kernel_Name", "Workgroup_Size", "Counter_Name",
                "Counter_Value"]
TRACE_COLS = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]
COPY = "__amd_rocclr_copyBuffer"


def write_csv(path, cols, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=cols)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def counters(d, name, rows):
    """rows: (kernel, grid, value) in dispatch order."""
    d.mkdir(exist_ok=True)
    write_csv(d / "run_counter_collection.csv", COUNTER_COLS,
              [{"Correlation_Id": i + 1, "Dispatch_Id": i + 1, "Grid_Size": g, "Kernel_Name": k,
                "Workgroup_Size": 256, "Counter_Name": name, "Counter_Value": f"{v:.6f}"}
               for i, (k, g, v) in enumerate(rows)])


def trace(d, launches):
    """launches: (kernel, duration_ns, gap_ns before the next one)."""
    d.mkdir(exist_ok=True)
    rows, t = [], 1_000_000
    for i, (k, dur, gap) in enumerate(launches):
        rows.append({"Dispatch_Id": i + 1, "Kernel_Name": k, "Start_Timestamp": t, "End_Timestamp": t + dur})
        t += dur + gap
    write_csv(d / "run_kernel_trace.csv", TRACE_COLS, rows)


def test_short_names():
    assert ps.short("void (anonymous namespace)::k_step_a<512, 3, true>(slam::PairDesc const*, int)") == \
        "k_step_a<512, 3, true>"
    assert ps.short("k_bilateral(unsigned short const*, int)") == "k_bilateral"


def test_stats_table(tmp_path):
    write_csv(tmp_path / "run_kernel_stats.csv",
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "MinNs", "MaxNs", "Percentage"],
              [{"Name": "void k_pyramid(PyrGrid)", "Calls": 3, "TotalDurationNs": 31500, "AverageNs": 10500,
                "MinNs": 9000, "MaxNs": 12000, "Percentage": 12.5}])
    out = ps.cmd_stats(str(tmp_path), 5, "python3 bench.py").splitlines()
    assert out[0] == "Command: `python3 bench.py`"
    assert out[-1] == "| `k_pyramid` | 3 | 0.03 | 10.50 | 9.00 | 12.00 | 12.50 |"


def test_kernel_traffic_scales_by_the_first_four_copies(tmp_path):
    """FETCH_SIZE reads 131072 KiB per 256 MiB copy (scale 2) and WRITE_SIZE
    262144 KiB (scale 1); the 1 GiB copies that follow on the same grid are
    not calibration copies."""
    fetch, write = tmp_path / "f", tmp_path / "w"
    counters(fetch, "FETCH_SIZE", [(COPY, 131072, 131072.0)] * 4 + [("k_splat_raster", 7813, 1000.0),
                                                                   ("k_splat_raster", 7813, 1200.0),
                                                                   (COPY, 131072, 524288.0)])
    counters(write, "WRITE_SIZE", [(COPY, 131072, 262144.0)] * 4 + [("k_splat_raster", 7813, 500.0),
                                                                   (COPY, 131072, 1048576.0)])
    r = ps.cmd_kernel_traffic(str(fetch), str(write), "k_splat_raster", 2_764_800)
    assert r["calibration"]["fetch_scale"] == 2.0 and r["calibration"]["write_scale"] == 1.0
    assert r["dispatches"] == [2, 1]
    assert r["fetch_bytes"] == 1100 * 2048 and r["write_bytes"] == 500 * 1024
    assert r["ratio"] == round((1100 * 2048 + 500 * 1024) / 2_764_800, 3)


def test_kernel_traffic_needs_a_calibration(tmp_path):
    fetch = tmp_path / "f"
    counters(fetch, "FETCH_SIZE", [("k_splat_raster", 7813, 1000.0)])
    with pytest.raises(ValueError):
        ps.cmd_kernel_traffic(str(fetch), str(fetch), "k_splat_raster", 1.0)


def test_batch_traffic_assigns_levels_in_schedule_order(tmp_path):
    """Two batched calls: each launches pass A and pass B 19 times, 4 at
    level 2, 5 at level 1, 10 at level 0.  Counter values of 10, 20, 30 KiB
    (A) and 1, 2, 3 KiB (B) at levels 2, 1, 0 come back per level."""
    a = "k_step_a<512, 2, false>"
    b = "k_step_b<512, 3>"
    val_a, val_b = {2: 10.0, 1: 20.0, 0: 30.0}, {2: 1.0, 1: 2.0, 0: 3.0}
    rows = []
    for _ in range(2):
        for lv in ps.CALL_LEVELS:
            rows += [(a, 4096, val_a[lv]), (b, 4096, val_b[lv])]
    fetch, write = tmp_path / "f", tmp_path / "w"
    counters(fetch, "FETCH_SIZE", [(COPY, 131072, 131072.0)] * 4 + rows)
    counters(write, "WRITE_SIZE", [(COPY, 131072, 262144.0)] * 4 + rows)
    r = ps.cmd_batch_traffic(str(fetch), str(write))
    for lv in range(3):
        la = r["levels"][str(lv)]["a"]
        assert la["dispatches"] == 2 * (4, 5, 10)[2 - lv]
        assert la["fetch_bytes"] == int(val_a[lv] * 2048) and la["write_bytes"] == int(val_a[lv] * 1024)
        joint = r["levels"][str(lv)]["joint"]
        assert joint["traffic_bytes"] == int(3072 * (val_a[lv] + val_b[lv]))
        assert joint["algorithmic_bytes"] == 62 * ps.BATCH_LEVEL_PX[lv]
    per_a = {lv: 3072 * val_a[lv] for lv in range(3)}
    assert r["a_traffic_bytes_per_launch_10_5_4"] == int((10 * per_a[0] + 5 * per_a[1] + 4 * per_a[2]) / 19)


def test_batch_traffic_rejects_a_partial_call(tmp_path):
    fetch = tmp_path / "f"
    counters(fetch, "FETCH_SIZE", [(COPY, 131072, 1.0)] * 4 + [("k_step_a<512, 2, false>", 4096, 1.0)] * 18)
    with pytest.raises(ValueError):
        ps.cmd_batch_traffic(str(fetch), str(fetch))


def test_runs_of_relaunches(tmp_path):
    """50 relaunches of 10 us, 2 us apart, then 50 of another kernel: each run
    reports its mean duration and start-to-start spacing; short runs and
    other kernels are left out."""
    a, b = "k_step_a<512, 3, true>(x)", "k_step_b<512, 3>(x)"
    trace(tmp_path, [("k_bilateral(x)", 5000, 1000)] + [(a, 10_000, 2000)] * 50 + [(b, 6000, 500)] * 50 +
          [(a, 10_000, 2000)] * 5)
    out = ps.cmd_runs(str(tmp_path), "k_step_", 40).splitlines()
    assert out[2:] == ["| `k_step_a<512, 3, true>` | 50 | 10.00 | 12.00 |",
                       "| `k_step_b<512, 3>` | 50 | 6.00 | 6.50 |"]


def frame(kernels, dur=5000, gap=1000):
    return [(k, dur, gap) for k in kernels] + [("k_track_end(x)", dur, gap)]


def test_frame_picks_plain_and_learned_calls(tmp_path):
    """Calls end at k_track_end; a plain call holds the SO3 loop, a learned
    one the seed policy, a batched one its own k_solve launches."""
    plain = ["k_bilateral", "k_so3_loop", "k_step_a<512, 3, true>", "k_step_b<512, 3>"]
    learned = ["k_bilateral", "k_seed_policy", "k_solve<512>", "k_step_a<512, 1, false>", "k_judge"]
    batched = ["k_so3_loop", "k_solve<512>", "k_step_a<512, 2, false>"]
    trace(tmp_path, [("k_track_end(x)", 5000, 1000)] + frame(plain) + frame(learned) + frame(batched) +
          frame(plain) + frame(plain))
    out = ps.cmd_frame(str(tmp_path), "plain")
    assert "k_seed_policy" not in out and "k_solve" not in out
    # five launches of 5 us, each after a 1 us gap: 30 us from the previous call's end
    assert out.endswith("One tracking call (frame) end to end: 30.0 us, 5 launches.")
    out = ps.cmd_frame(str(tmp_path), "learned")
    assert "| `k_seed_policy` | 1 | 5.0 | 5.00 |" in out and "6 launches" in out
