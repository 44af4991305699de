"""The data formats and trajectory evaluators around the odometry
(slam_amd/dataio.py, slam_amd/evaluate.py), pinned by closed forms: round
trips, rigid transforms that alignment must remove, constant errors of known
size, and scipy for the rotation conversions."""
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from slam_amd import dataio, evaluate, synth


def _random_pose(rng, t_scale=1.0):
    T = np.eye(4)
    T[:3, :3] = Rotation.random(random_state=rng.integers(1 << 30)).as_matrix()
    T[:3, 3] = rng.normal(size=3) * t_scale
    return T


# ------------------------------------------------------------------ poses

def test_quaternion_matches_scipy_and_eigen_sign():
    rng = np.random.default_rng(0)
    for _ in range(200):
        R = Rotation.random(random_state=rng.integers(1 << 30)).as_matrix()
        q = dataio.quaternion_from_matrix(R)
        ref = Rotation.from_matrix(R).as_quat()           # (x, y, z, w)
        assert min(np.abs(q - ref).max(), np.abs(q + ref).max()) < 1e-12
        # Eigen's branches: w > 0 when the trace is positive, else the
        # component of the largest diagonal entry is positive
        if np.trace(R) > 0:
            assert q[3] > 0
        else:
            assert q[int(np.argmax(np.diag(R)))] > 0
    # a rotation of pi about x (trace -1): the x branch
    q = dataio.quaternion_from_matrix(np.diag([1.0, -1.0, -1.0]))
    assert np.allclose(q, [1, 0, 0, 0])


def test_matrix_from_pose_inverts_quaternion_and_normalises():
    rng = np.random.default_rng(1)
    for _ in range(50):
        T = _random_pose(rng)
        q = dataio.quaternion_from_matrix(T[:3, :3])
        assert np.allclose(dataio.matrix_from_pose(T[:3, 3], q), T, atol=1e-12)
        # a scaled quaternion is the same rotation
        assert np.allclose(dataio.matrix_from_pose(T[:3, 3], 3.7 * q), T, atol=1e-12)
    assert np.array_equal(dataio.matrix_from_pose([1, 2, 3], [0, 0, 0, 0])[:3, :3], np.eye(3))


def test_pose_log_format_and_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    poses = [_random_pose(rng).astype(np.float32) for _ in range(20)]
    ticks = list(range(1, 21))
    path = str(tmp_path / "poses.txt")
    dataio.write_pose_log(path, ticks, poses)
    lines = open(path).read().splitlines()
    assert len(lines) == 20
    fields = lines[0].split(" ")
    assert len(fields) == 8 and fields[0] == "1"
    # six significant digits, as the C++ stream default
    for f in fields[1:]:
        assert len(f.lstrip("-").replace(".", "").lstrip("0").split("e")[0]) <= 6
    traj = dataio.read_trajectory(path)
    for t, T in zip(ticks, poses):
        assert np.allclose(traj[float(t)], T, atol=2e-5)


def test_read_trajectory_skips_comments_zero_quaternions_and_nans(tmp_path):
    path = tmp_path / "t.txt"
    path.write_text("# header\n1 0 0 0 0 0 0 1\n2,1,2,3,0,0,0,0\n3\t1\t1\t1\t0\t0\t0\t1\n4 nan 0 0 0 0 0 1\n")
    traj = dataio.read_trajectory(str(path))
    assert sorted(traj) == [1.0, 3.0]
    assert np.allclose(traj[3.0][:3, 3], [1, 1, 1])


# ------------------------------------------------------------- TUM lists

def test_associate_greedy_unique_matching():
    first = {1.00: ["a"], 1.10: ["b"], 1.20: ["c"]}
    second = {1.005: ["x"], 1.012: ["y"], 1.19: ["z"], 5.0: ["w"]}
    # 1.00 is closest to 1.005 (0.005) and to 1.012 (0.012): the smaller
    # difference wins and 1.012 stays unmatched; 1.10 has nothing within 0.02
    assert dataio.associate(first, second, 0.0, 0.02) == [(1.00, 1.005), (1.20, 1.19)]
    # an offset shifts the second stamps
    assert dataio.associate({2.0: []}, {1.0: []}, 1.0, 0.02) == [(2.0, 1.0)]


def test_depth_png_scaling_rounds_to_nearest():
    raw = np.array([0, 1, 2, 3, 5, 7, 8, 12, 13, 5000, 65535], dtype=np.uint16)
    # x * 0.2: 0, .2, .4, .6, 1, 1.4, 1.6, 2.4, 2.6, 1000, 13107
    assert dataio.depth_png_to_mm(raw).tolist() == [0, 0, 0, 1, 1, 1, 2, 2, 3, 1000, 13107]


def test_tum_frames_read_lists_in_step(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(3)
    h, w = 12, 16
    rgb_lines, depth_lines, frames = [], [], []
    for k in range(4):
        rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        depth = rng.integers(0, 30000, (h, w), dtype=np.uint16)
        Image.fromarray(rgb).save(tmp_path / f"rgb{k}.png")
        Image.fromarray(depth).save(tmp_path / f"depth{k}.png")
        rgb_lines.append(f"{k * 0.033:.6f} rgb{k}.png")
        depth_lines.append(f"{k * 0.033 + 0.001:.6f} depth{k}.png")
        frames.append((rgb, depth))
    (tmp_path / "rgb.txt").write_text("\n".join(rgb_lines) + "\n")
    (tmp_path / "depth.txt").write_text("\n".join(depth_lines) + "\n")
    seq = dataio.TumFrames(str(tmp_path / "rgb.txt"), str(tmp_path / "depth.txt"), str(tmp_path))
    got = list(seq)
    assert len(got) == 3                      # the last listed frame is not read
    for k, fr in enumerate(got):
        assert np.array_equal(fr.rgb, frames[k][0][..., ::-1])      # blue, green, red
        assert np.array_equal(fr.depth, dataio.depth_png_to_mm(frames[k][1]))


# ------------------------------------------------------------------- .klg

@pytest.mark.parametrize("compress,image", [(True, "raw"), (False, "raw"), (True, "none"), (True, "jpeg")])
def test_klg_round_trip(tmp_path, compress, image):
    rng = np.random.default_rng(4)
    h, w = 24, 32
    frames = [dataio.Frame(1000 + k, rng.integers(0, 8000, (h, w), dtype=np.uint16),
                           rng.integers(0, 256, (h, w, 3), dtype=np.uint8)) for k in range(3)]
    if image == "jpeg":
        # smooth colour so the lossy round trip is close
        yy, xx = np.mgrid[0:h, 0:w]
        for fr in frames:
            fr.rgb[:] = np.stack([xx * 7, yy * 9, (xx + yy) * 4], -1).astype(np.uint8)
    path = str(tmp_path / "seq.klg")
    dataio.write_klg(path, frames, compress_depth=compress, image=image)
    got = list(dataio.read_klg(path, w, h))
    assert len(got) == 3
    for a, b in zip(frames, got):
        assert b.timestamp == a.timestamp
        assert np.array_equal(b.depth, a.depth)
        if image == "raw":
            assert np.array_equal(b.rgb, a.rgb)
        elif image == "none":
            assert not b.rgb.any()
        else:
            assert np.abs(b.rgb.astype(int) - a.rgb.astype(int)).mean() < 3
    flipped = next(dataio.read_klg(path, w, h, flip_colors=True))
    assert np.array_equal(flipped.rgb, got[0].rgb[..., ::-1])


# ------------------------------------------------------------- evaluators

def _traj(poses, dt=1 / 30):
    return evaluate.trajectory([i * dt for i in range(len(poses))], poses)


def test_ate_removes_a_rigid_transform():
    poses = synth.spline_trajectory(40, seed=5, max_step_m=0.05, max_step_deg=3.0)
    rng = np.random.default_rng(6)
    G = _random_pose(rng, 2.0)
    moved = [G @ T for T in poses]
    r = evaluate.ate(_traj(poses), _traj(moved))
    assert r["rmse"] < 1e-9 and r["pairs"] == 40
    # the alignment found is G^-1 (est -> gt)
    Gi = np.linalg.inv(G)
    assert np.allclose(r["rotation"], Gi[:3, :3], atol=1e-9)
    assert np.allclose(r["translation"], Gi[:3, 3], atol=1e-9)


def test_ate_of_a_known_perpendicular_offset():
    # ground truth in the plane z = 0, every (x, y) twice; the estimate puts
    # one copy at z = +d and the other at z = -d.  The offsets have zero mean
    # and no correlation with x or y, so the best rigid fit is the identity
    # and every error after alignment is d.
    d = 0.01
    xy = [(0.1 * i, 0.05 * (i % 3)) for i in range(10)]
    gt, est = [], []
    for sign in (1, -1):
        for x, y in xy:
            g = np.eye(4)
            g[0, 3], g[1, 3] = x, y
            e = g.copy()
            e[2, 3] = sign * d
            gt.append(g)
            est.append(e)
    r = evaluate.ate(_traj(gt), _traj(est))
    assert abs(r["rmse"] - d) < 1e-12 and abs(r["max"] - d) < 1e-12 and abs(r["min"] - d) < 1e-12


def test_rpe_zero_and_constant_step_error():
    poses = synth.spline_trajectory(30, seed=8, max_step_m=0.02, max_step_deg=1.0)
    r = evaluate.rpe(_traj(poses), _traj(poses), delta=1, delta_unit="f")
    # arccos near 1 turns rounding of the trace into ~1e-6 deg
    assert r["translational"]["max"] < 1e-12 and r["rotational_deg"]["max"] < 1e-5
    # every estimated step carries an extra motion E: est_{k+1} = est_k E M_k
    # with M_k = gt_k^-1 gt_{k+1}; the relative motions are then
    # est_{k+1}^-1 est_k = M_k^-1 E^-1 and gt_{k+1}^-1 gt_k = M_k^-1, and the
    # error (M_k^-1 E^-1)^-1 M_k^-1 = E for every pair
    E = synth.make_pose(synth.so3_exp(np.array([0.0, 0.0, np.radians(0.5)])), np.array([0.003, 0.0, 0.004]))
    est = [poses[0]]
    for k in range(len(poses) - 1):
        M = np.linalg.inv(poses[k]) @ poses[k + 1]
        est.append(est[-1] @ E @ M)
    r = evaluate.rpe(_traj(poses), _traj(est), delta=1, delta_unit="f")
    assert abs(r["translational"]["mean"] - 0.005) < 1e-9
    assert abs(r["rotational_deg"]["mean"] - 0.5) < 1e-6
    # a two-frame delta compounds two steps
    r2 = evaluate.rpe(_traj(poses), _traj(est), delta=2, delta_unit="f")
    assert r2["translational"]["mean"] > r["translational"]["mean"]


def test_closest_index_matches_brute_force():
    rng = np.random.default_rng(9)
    L = sorted(rng.uniform(0, 10, 50).tolist())
    for t in rng.uniform(-1, 11, 200):
        i = evaluate.closest_index(L, t)
        assert abs(L[i] - t) == min(abs(v - t) for v in L)
