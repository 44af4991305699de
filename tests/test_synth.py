"""The seeded renderer that makes every workload: deterministic, in the
configured depth range, with the configured dropout, and its noise-free model
prediction consistent with the sensor depth."""
from __future__ import annotations

import numpy as np

from slam_amd import synth


def test_render_is_deterministic_and_in_range():
    cam = synth.CAM_320
    a = synth.Scene(5, 20).render(synth.base_pose(), cam, noise_seed=9)
    b = synth.Scene(5, 20).render(synth.base_pose(), cam, noise_seed=9)
    assert np.array_equal(a.depth, b.depth) and np.array_equal(a.rgb, b.rgb)
    assert np.array_equal(a.vert4, b.vert4)
    d = a.depth
    valid = d > 0
    assert d[valid].min() >= 500 and d[valid].max() <= 5000
    hole = 1.0 - valid.mean()
    assert 0.01 < hole < 0.2
    c = synth.Scene(6, 20).render(synth.base_pose(), cam, noise_seed=9)
    assert not np.array_equal(a.depth, c.depth)


def test_model_prediction_matches_sensor_depth():
    cam = synth.CAM_320
    f = synth.Scene(7, 20).render(synth.base_pose(), cam, noise_seed=1)
    z = f.vert4[..., 2]
    valid = (f.depth > 0) & (z > 0)
    err = np.abs(f.depth[valid] / 1000.0 - z[valid])
    # noise sigma is 1 mm * z^2 (z <= 5 m): allow 6 sigma plus quantisation
    assert np.all(err <= 6e-3 * z[valid] ** 2 + 1e-3)
    n = f.norm4[..., :3][valid]
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
    # normals face the camera
    v = f.vert4[..., :3][valid]
    assert np.mean(np.sum(n * v, axis=1) < 0) > 0.99


def test_perturb_bounds():
    rng = np.random.default_rng(0)
    T0 = synth.base_pose()
    for _ in range(50):
        T1 = synth.perturb(T0, rng, 0.03, 2.0)
        dt, dr = synth.pose_error(T1, T0)
        assert dt <= 0.03 + 1e-12 and dr <= 2.0 + 1e-9


def test_spline_step_bounds():
    poses = synth.spline_trajectory(100, 7)
    steps = [synth.pose_error(poses[i + 1], poses[i]) for i in range(99)]
    assert max(s[0] for s in steps) <= 0.01 + 1e-9
    assert max(s[1] for s in steps) <= 0.5 + 1e-6


def test_long_burst_sequence_reaches_its_step_bound_inside_the_room():
    """The fast-motion stress sequence (2000 frames, bursts up to 10 cm /
    8 deg per frame) realises its bound while the camera stays within the
    0.8 m excursion: the path folds back rather than its steps shrinking."""
    poses = synth.spline_trajectory(2000, 920, max_step_m=0.10, max_step_deg=8.0, bursts=True)
    steps = [synth.pose_error(poses[i + 1], poses[i]) for i in range(len(poses) - 1)]
    dt = max(s[0] for s in steps)
    dr = max(s[1] for s in steps)
    assert 0.9 * 0.10 <= dt <= 0.10 + 1e-9
    assert 0.9 * 8.0 <= dr <= 8.0 + 1e-6
    ext = max(np.abs(T[:3, 3] - poses[0][:3, 3]).max() for T in poses)
    assert ext <= 0.8 + 1e-9
    # most frames move at the base speed; the bursts carry the largest steps
    assert np.median([s[0] for s in steps]) < 0.5 * dt
