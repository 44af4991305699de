"""Model-to-model registration and the loop-closure acceptance test of the
reference's frame loop (elastic_fusion_file.cpp:449-498).

When the loop closes, the inactive part of the model and the active part are
both predicted from the current pose; a second odometry object registers the
one to the other with ICP and RGB and SO3 off, and the result is accepted
when every diagonal entry of the covariance (inverse of the last 6x6 system)
is at most ``covThresh``, more than ``icpCountThresh`` pixels were ICP
inliers and the ICP error is below ``icpErrThresh``.  The thresholds default
to the values of the reference's configs (src/configs/ef_file.cfg:9-14).

The registration and the test run on the device through the C ABI
(odom_loop_check); this module only orders the calls as the frame loop does.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .odometry import RGBDOdometry

COV_THRESH = 1e-5          # ef_file.cfg:10
ICP_ERR_THRESH = 5e-5      # ef_file.cfg:9
ICP_COUNT_THRESH = 40000   # ef_file.cfg:14


@dataclass
class LoopResult:
    accepted: bool
    pose: np.ndarray           # 4x4 float32: the registered estimate (estPose)
    stats: dict                # lastICPError, lastICPCount, ... of the registration


def model_to_model(odo: RGBDOdometry, inactive_vertices, inactive_normals, inactive_image,
                   active_vertices, active_normals, active_image, pose, depthCutoff: float,
                   pyramid: bool = True, fastOdom: bool = False, covThresh: float = COV_THRESH,
                   icpCountThresh: float = ICP_COUNT_THRESH, icpErrThresh: float = ICP_ERR_THRESH) -> LoopResult:
    """Register the active model prediction to the inactive one from ``pose``
    (4x4) and apply the acceptance test (elastic_fusion_file.cpp:460-498).

    The vertex/normal maps are (h, w, 4) float32 in the camera frame of
    ``pose``; the images (h, w, 3 or 4) uint8.  ``odo`` is the registration's
    own odometry object (the frame loop's ``modelToModel``), not the
    frame-to-model tracker."""
    pose = np.asarray(pose, dtype=np.float32)
    # initICP* before initRGB*: the RGB inits read the vertex maps
    odo.initICPModel(inactive_vertices, inactive_normals, depthCutoff, pose)
    odo.initRGBModel(inactive_image)
    odo.initICPMaps(active_vertices, active_normals, depthCutoff)
    odo.initRGB(active_image)
    t, R = odo.getIncrementalTransformation(pose[:3, 3], pose[:3, :3], False, 10.0, pyramid, fastOdom, False)
    accepted = odo.loopCheck(covThresh, icpCountThresh, icpErrThresh)
    est = np.eye(4, dtype=np.float32)
    est[:3, :3] = np.asarray(R, np.float32).reshape(3, 3)
    est[:3, 3] = np.asarray(t, np.float32).reshape(3)
    return LoopResult(accepted, est, odo.stats())
