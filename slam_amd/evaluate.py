"""Trajectory evaluation as the reference's benchmark tools define it.

- ATE (``benchmark/evaluate_ate.py:47-79, 162``): associate the estimated to
  the ground-truth stamps (associate.py), align the estimated positions to
  the ground truth with Horn's closed form (SVD of the cross-covariance, with
  the reflection fixed), and report the statistics of the per-pose
  translational errors after alignment.
- RPE (``benchmark/evaluate_rpe.py:204-298, 367``): for pairs of estimated
  poses a fixed delta apart (frames, seconds, metres or radians along the
  estimated trajectory), the error of the relative motion
  ominus(ominus(est_j, est_i), ominus(gt_j, gt_i)) with ominus(a, b) = a^-1 b,
  as translational distance and rotation angle.  Only the fixed-delta form
  is provided (the reference's all-pairs form samples pairs at random).

Trajectories are {stamp: 4x4 pose} dicts (``dataio.read_trajectory``).
"""
from __future__ import annotations

import numpy as np

from .dataio import associate


def horn_align(model: np.ndarray, data: np.ndarray):
    """Rotation R and translation t minimising |R model + t - data| for 3xN
    point sets; also the per-point errors after alignment."""
    mc = model - model.mean(1, keepdims=True)
    dc = data - data.mean(1, keepdims=True)
    W = mc @ dc.T                                  # sum of outer(model_i, data_i)
    U, _, Vh = np.linalg.svd(W.T)
    S = np.eye(3)
    if np.linalg.det(U) * np.linalg.det(Vh) < 0:
        S[2, 2] = -1
    R = U @ S @ Vh
    t = data.mean(1, keepdims=True) - R @ model.mean(1, keepdims=True)
    err = np.sqrt(np.sum((R @ model + t - data) ** 2, axis=0))
    return R, t, err


def _stats(e: np.ndarray) -> dict:
    e = np.asarray(e, dtype=np.float64)
    return {"rmse": float(np.sqrt(e @ e / len(e))), "mean": float(e.mean()), "median": float(np.median(e)),
            "std": float(e.std()), "min": float(e.min()), "max": float(e.max()), "pairs": int(len(e))}


def ate(gt: dict, est: dict, offset: float = 0.0, max_difference: float = 0.02, scale: float = 1.0) -> dict:
    """Absolute trajectory error (metres) of est against gt."""
    matches = associate(gt, est, offset, max_difference)
    if len(matches) < 2:
        raise ValueError("fewer than two matching stamps between the trajectories")
    first = np.array([gt[a][:3, 3] for a, _ in matches]).T
    second = np.array([est[b][:3, 3] * scale for _, b in matches]).T
    R, t, err = horn_align(second, first)
    out = _stats(err)
    out["rotation"] = R
    out["translation"] = t.ravel()
    return out


def ominus(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """The motion from pose a to pose b: a^-1 b."""
    return np.linalg.inv(a) @ b


def motion_distance(T: np.ndarray) -> float:
    return float(np.linalg.norm(T[:3, 3]))


def motion_angle(T: np.ndarray) -> float:
    return float(np.arccos(min(1.0, max(-1.0, (np.trace(T[:3, :3]) - 1) / 2))))


def closest_index(L, t) -> int:
    """Index of the value of the sorted list L closest to t (a bisection that
    keeps the best seen, as evaluate_rpe.py find_closest_index)."""
    lo, hi = 0, len(L)
    best, diff = 0, abs(L[0] - t)
    while lo < hi:
        mid = (lo + hi) // 2
        if abs(L[mid] - t) < diff:
            diff = abs(L[mid] - t)
            best = mid
        if t == L[mid]:
            return mid
        if L[mid] > t:
            hi = mid
        else:
            lo = mid + 1
    return best


def rpe(gt: dict, est: dict, delta: float = 1.0, delta_unit: str = "f", offset: float = 0.0,
        scale: float = 1.0) -> dict:
    """Relative pose error (metres, degrees) over pairs of estimated poses a
    fixed `delta` apart in `delta_unit` ('f' frames, 's' seconds, 'm' metres
    or 'rad' / 'deg' of rotation along the estimated trajectory)."""
    sg, se = sorted(gt), sorted(est)
    if delta_unit == "s":
        index = list(se)
    elif delta_unit == "f":
        index = list(range(len(se)))
    elif delta_unit in ("m", "rad", "deg"):
        f = motion_distance if delta_unit == "m" else motion_angle
        k = 180.0 / np.pi if delta_unit == "deg" else 1.0
        index = [0.0]
        for i in range(len(se) - 1):
            index.append(index[-1] + f(ominus(est[se[i + 1]], est[se[i]])) * k)
    else:
        raise ValueError(f"unknown delta unit {delta_unit!r}")
    pairs = []
    for i in range(len(se)):
        j = closest_index(index, index[i] + delta)
        if j != len(se) - 1:
            pairs.append((i, j))
    gt_dt = float(np.median(np.diff(sg))) if len(sg) > 1 else 0.0
    trans, rot = [], []
    for i, j in pairs:
        g0 = sg[closest_index(sg, se[i] + offset)]
        g1 = sg[closest_index(sg, se[j] + offset)]
        if abs(g0 - (se[i] + offset)) > 2 * gt_dt or abs(g1 - (se[j] + offset)) > 2 * gt_dt:
            continue
        rel_e = ominus(est[se[j]], est[se[i]])
        rel_e[:3, 3] *= scale
        e = ominus(rel_e, ominus(gt[g1], gt[g0]))
        trans.append(motion_distance(e))
        rot.append(motion_angle(e))
    if len(trans) < 2:
        raise ValueError("fewer than two pose pairs to compare")
    out = {"translational": _stats(trans)}
    out["rotational_deg"] = _stats(np.degrees(rot))
    return out


def trajectory(stamps, poses) -> dict:
    """{stamp: 4x4} from parallel sequences."""
    return {float(s): np.asarray(T, dtype=np.float64) for s, T in zip(stamps, poses)}
