"""The reference's frame loop without loop closure, on the device: bilateral
filter, frame-to-model tracking against the filled-in splat prediction, the
model update, and the next prediction (elastic_fusion_file.cpp:335-420,
551-602, with closeLoops off).

Every stage runs through the C ABI on one HIP stream (the odometry's); the
host reads back only what the reference's loop reads: the tracked pose (for
the fusion weighting and the next prediction) and the model's surfel count.
The prediction that the reference makes right after tracking
(elastic_fusion_file.cpp:399) feeds only the loop-closure search, so it is
not made here; the one at the end of the frame (:601) is.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import splat, surfels
from .odometry import FMT_DEPTH_U16, MEM_DEVICE, RGBDOdometry, _check, frame_view, lib as odom_lib


@dataclass
class LoopParams:
    """The reference configs' values (src/configs/ef_file.cfg)."""
    depth_cutoff: float = 3.0        # depthCutoff: bilateral and metric depth window (m)
    max_depth: float = 20.0          # maxDepthProcessed
    confidence: float = 10.0
    time_delta: int = 200
    icp_weight: float = 10.0
    so3: bool = True
    max_surfels: int = 4_000_000
    learned: bool = False            # seed each call from the warm-start network (no SO3 pass then)
    learned_iters: tuple = (10, 5, 4)


def _rodrigues_norm(R: np.ndarray) -> float:
    c = (np.trace(R) - 1.0) / 2.0
    return float(np.arccos(np.clip(c, -1.0, 1.0)))


class FrameLoop:
    """Processes frames (uint16 depth in mm, RGB8) held in device memory."""

    def __init__(self, width: int, height: int, cx: float, cy: float, fx: float, fy: float,
                 params: LoopParams | None = None):
        import torch
        self.p = params or LoopParams()
        self.w, self.h = width, height
        self.odo = RGBDOdometry(width, height, cx, cy, fx, fy)
        stream = self.odo.stream
        self.pred = splat.IndexMap(width, height, cx, cy, fx, fy)
        self.pred.set_stream(stream)
        self.model = surfels.SurfelModel(width, height, cx, cy, fx, fy, self.p.max_surfels)
        self.model.set_stream(stream)
        self.filtered = torch.empty((height, width), dtype=torch.uint16, device="cuda")
        self.pose = np.eye(4, dtype=np.float32)
        self.odo.set_pose(self.pose[:3, 3], self.pose[:3, :3])
        self.net = None
        if self.p.learned:
            from . import warmstart as ws
            self.ws = ws
            self.net = ws.WarmStartNet(max_batch=1)
            self.net.load()
            self.net.set_stream(stream)
            self.x = torch.zeros((1, 60, 80, 4), dtype=torch.bfloat16, device="cuda")
            self.pred6 = torch.zeros((1, 6), dtype=torch.float32, device="cuda")
        self.tick = 0
        self.last_icp_error = 0.0
        self.poses: list[np.ndarray] = []

    def _bilateral(self, depth):
        v, keep = frame_view(depth, FMT_DEPTH_U16)
        _check("odom_bilateral_depth", odom_lib().odom_bilateral_depth(
            self.odo.handle, ctypes.byref(v), self.p.depth_cutoff, ctypes.c_void_p(self.filtered.data_ptr()),
            MEM_DEVICE))
        return self.filtered

    def _predict(self):
        """predict() (elastic_fusion_file.cpp:17-44): the active splat at the
        current pose, then the fill-in from the filtered depth and colour."""
        self.pred.combinedPredict(self.pose, self.model.device_model(), self.p.max_depth, self.p.confidence,
                                  self.tick, self.tick, self.p.time_delta)

    def process(self, rgb, depth) -> np.ndarray:
        """One frame; returns the tracked camera-to-model pose (4x4)."""
        self.tick += 1
        p = self.p
        filtered = self._bilateral(depth)
        if self.tick == 1:
            self.model.initialise(rgb, depth, filtered, p.depth_cutoff, self.tick, p.max_depth)
            self.odo.initFirstRGB(rgb)
        else:
            last = self.pose.copy()
            tracking_ok = self.last_icp_error < 1e-4      # the previous call's error, as the loop reads it
            self.odo.initICPModel(self.pred.device(splat.FILL_VERTEX), self.pred.device(splat.FILL_NORMAL),
                                  p.max_depth, self.pose)
            self.odo.initRGBModel(self.pred.device(splat.FILL_IMAGE))
            self.odo.initICP(filtered, p.max_depth)
            self.odo.initRGB(rgb)
            if self.net is not None:
                # the learned warm start: the network's motion seeds the call,
                # which keeps it only if its association is at least as good
                self.odo.initWarmStartInput(self.x)
                self.ws._lib().cnn_forward(self.net.handle, self.x.data_ptr(), 1, self.pred6.data_ptr())
                self.odo.track_seeded(self.pred6, p.learned_iters, False, p.icp_weight, True, self.ws.T_SCALE,
                                      self.ws.R_SCALE)
                t, R = self.odo.get_pose()
            else:
                t, R = self.odo.getIncrementalTransformation(self.pose[:3, 3], self.pose[:3, :3], False,
                                                             p.icp_weight, True, False, p.so3)
            self.pose = np.eye(4, dtype=np.float32)
            self.pose[:3, :3] = np.asarray(R, np.float32).reshape(3, 3)
            self.pose[:3, 3] = np.asarray(t, np.float32).reshape(3)
            self.last_icp_error = self.odo.stats()["lastICPError"]
            # weight by velocity (elastic_fusion_file.cpp:377-398)
            diff = np.linalg.inv(self.pose.astype(np.float64)) @ last.astype(np.float64)
            weighting = max(float(np.linalg.norm(diff[:3, 3])), _rodrigues_norm(diff[:3, :3]))
            weighting = max(1.0 - min(weighting, 0.01) / 0.01, 0.5)
            if tracking_ok:
                self.model.predictIndices(self.pose, self.tick, p.max_depth, p.time_delta)
                self.model.fuse(self.pose, self.tick, rgb, depth, filtered, p.depth_cutoff, p.max_depth, weighting)
                self.model.predictIndices(self.pose, self.tick, p.max_depth, p.time_delta)
                self.model.clean(self.pose, self.tick, p.confidence, p.time_delta)
        self._predict()
        self.pred.fillIn(filtered, rgb)
        self.poses.append(self.pose.copy())
        return self.pose
