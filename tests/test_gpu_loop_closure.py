"""Model-to-model registration and the loop-closure acceptance test on the
device (odom_loop_check) against the oracle: the registered pose, the
covariance diagonal and the accept/reject decision for thresholds on both
sides of the measured values, and a degenerate registration (no valid active
vertex) that must be rejected without an error."""
import numpy as np
import pytest

from slam_amd import synth
from slam_amd.odometry import COV_THRESH, ICP_COUNT_THRESH, ICP_ERR_THRESH, RGBDOdometry

from helpers import CUTOFF, make_pair, pose_of

pytestmark = pytest.mark.gpu


def mk(cam):
    return RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)


def register(odo, f0, f1, T0, active_vert4=None):
    """The frame loop's model-to-model calls (elastic_fusion_file.cpp:460-474):
    the inactive prediction f0 and the active prediction f1, both as maps."""
    v = f1.vert4 if active_vert4 is None else active_vert4
    odo.initICPModel(f0.vert4, f0.norm4, CUTOFF, T0.astype(np.float32))
    odo.initRGBModel(f0.rgba)
    odo.initICPMaps(v, f1.norm4, CUTOFF)
    odo.initRGB(f1.rgba)
    t, R = odo.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 10.0, True, False, False)
    return pose_of(t, R)


@pytest.mark.parametrize("cam,seed", [(synth.CAM_320, 31), (synth.CAM_640, 32)], ids=["320x240", "640x480"])
def test_loop_check_matches_oracle(orc, cam, seed):
    T0, T1, f0, f1 = make_pair(cam, seed)
    gpu = mk(cam)
    ref = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    Tg, Tr = register(gpu, f0, f1, T0), register(ref, f0, f1, T0)
    dt, dr = synth.pose_error(Tg, Tr)
    assert dt <= 1e-3 and dr <= 0.1, (dt, dr)
    dt, dr = synth.pose_error(Tg, T1)
    assert dt < 5e-3 and dr < 0.3, (dt, dr)

    cg, cr = np.diag(gpu.getCovariance()), np.diag(ref.getCovariance())
    assert np.allclose(cg, cr, rtol=1e-2), (cg, cr)
    s = ref.stats()
    cov, cnt, err = float(cr.max()), s["lastICPCount"], s["lastICPError"]
    assert cnt > 0 and err > 0
    decisions = []
    for fc in (0.5, 2.0):
        for fn in (0.5, 1.5):
            for fe in (0.5, 2.0):
                th = (cov * fc, cnt * fn, err * fe)
                want = ref.loop_accept(*th)
                assert gpu.loopCheck(*th) == want, th
                decisions.append(want)
    assert any(decisions) and not all(decisions)
    # exactly at a threshold: <= for the covariance, strict for count and error
    assert gpu.loopCheck(float(cg.max()), 0.0, 1.0)
    assert not gpu.loopCheck(1.0, gpu.stats()["lastICPCount"], 1.0)
    assert not gpu.loopCheck(1.0, 0.0, gpu.stats()["lastICPError"])


@pytest.mark.parametrize("cam,seed", [(synth.CAM_320, 33), (synth.CAM_640, 34)], ids=["320x240", "640x480"])
def test_model_to_model_default_thresholds(orc, cam, seed):
    """RGBDOdometry.modelToModel with the reference configs' thresholds
    decides as the oracle does (the case is skipped where a measured value sits within 5% of
    a threshold, where FP32 summation order could flip it)."""
    T0, T1, f0, f1 = make_pair(cam, seed)
    accepted, est = mk(cam).modelToModel(f0.vert4, f0.norm4, f0.rgba, f1.vert4, f1.norm4, f1.rgba, T0, CUTOFF)
    ref = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    Tr = register(ref, f0, f1, T0)
    dt, dr = synth.pose_error(est.astype(np.float64), Tr)
    assert dt <= 1e-3 and dr <= 0.1, (dt, dr)
    s = ref.stats()
    cov = float(np.diag(ref.getCovariance()).max())
    near = [abs(v / th - 1) < 0.05 for v, th in ((cov, COV_THRESH), (s["lastICPCount"], ICP_COUNT_THRESH),
                                                 (s["lastICPError"], ICP_ERR_THRESH))]
    if any(near):
        pytest.skip(f"measured values within 5% of a threshold: cov {cov} stats {s}")
    want = ref.loop_accept(COV_THRESH, ICP_COUNT_THRESH, ICP_ERR_THRESH)
    assert accepted == want, (cov, s)


def test_degenerate_registration_is_rejected(orc):
    """No valid active vertex: zero information, a singular last system, no
    covariance -- rejected, as the oracle decides, and no error raised."""
    cam = synth.CAM_320
    T0, T1, f0, f1 = make_pair(cam, 35)
    nan = np.full_like(f1.vert4, np.nan)
    gpu = mk(cam)
    ref = orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    register(gpu, f0, f1, T0, nan)
    register(ref, f0, f1, T0, nan)
    assert gpu.stats()["lastICPCount"] == 0
    assert not ref.loop_accept(1e30, -1.0, 1e30)
    assert not gpu.loopCheck(1e30, -1.0, 1e30)
