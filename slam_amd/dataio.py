"""Data formats on either side of the odometry: the pose log the frame loop
writes, the TUM RGB-D frame lists and the ``.klg`` raw log it reads.

- Pose log (``elastic_fusion_file.cpp:383-387``): one line per tracked frame,
  ``tick tx ty tz qx qy qz qw``, written with the C++ stream defaults (six
  significant digits) and the quaternion of ``Eigen::Quaternionf(rot)``.
- TUM lists (``FileReader.cpp:20-57``): ``rgb.txt`` / ``depth.txt`` read line
  by line in step (the lists are associated beforehand); colour images in
  cv::imread's (blue, green, red) byte order; depth PNGs are
  16-bit in units of 1/5000 m and converted to millimetres by
  ``convertTo(CV_16UC1, 1000.0 / 5000)``, i.e. scaled in double precision and
  rounded to the nearest integer (ties to even), saturated to uint16.
- ``.klg`` (``RawLogReader.cpp:20-130``): int32 frame count, then per frame
  int64 timestamp, int32 depth size, int32 image size, the depth (raw
  uint16 when its size is 2 * w * h, zlib otherwise), the image (raw RGB
  when 3 * w * h bytes, JPEG when smaller and non-zero, black when 0).
"""
from __future__ import annotations

import io
import os
import struct
import zlib
from dataclasses import dataclass
from typing import Iterator

import numpy as np

# ------------------------------------------------------------------ poses


def quaternion_from_matrix(R: np.ndarray) -> np.ndarray:
    """(x, y, z, w) of a rotation matrix, with Eigen's choice of branch and
    sign (Eigen/src/Geometry/Quaternion.h, quaternionbase_assign_impl): if
    the trace is positive w = sqrt(1 + trace) / 2 > 0, else the largest
    diagonal entry picks the component computed first, which is positive."""
    m = np.asarray(R)
    if m.dtype not in (np.float32, np.float64):
        m = m.astype(np.float64)
    one, half = m.dtype.type(1.0), m.dtype.type(0.5)
    t = m[0, 0] + m[1, 1] + m[2, 2]
    q = np.zeros(4, dtype=m.dtype)
    if t > 0:
        s = np.sqrt(t + one)
        w = half * s
        s = half / s
        q[:] = ((m[2, 1] - m[1, 2]) * s, (m[0, 2] - m[2, 0]) * s, (m[1, 0] - m[0, 1]) * s, w)
    else:
        i = 0
        if m[1, 1] > m[0, 0]:
            i = 1
        if m[2, 2] > m[i, i]:
            i = 2
        j, k = (i + 1) % 3, (i + 2) % 3
        s = np.sqrt(m[i, i] - m[j, j] - m[k, k] + one)
        q[i] = half * s
        s = half / s
        q[3] = (m[k, j] - m[j, k]) * s
        q[j] = (m[j, i] + m[i, j]) * s
        q[k] = (m[k, i] + m[i, k]) * s
    return q


def matrix_from_pose(t, q) -> np.ndarray:
    """4x4 pose from a translation and a quaternion (x, y, z, w), not
    necessarily unit: scaled by sqrt(2 / |q|^2) before the outer product
    (evaluate_rpe.py transform44); a zero quaternion gives the identity."""
    T = np.eye(4)
    T[:3, 3] = np.asarray(t, dtype=np.float64)
    q = np.array(q, dtype=np.float64)
    n = float(q @ q)
    if n < np.finfo(float).eps * 4.0:
        return T
    q *= np.sqrt(2.0 / n)
    o = np.outer(q, q)
    T[:3, :3] = [[1.0 - o[1, 1] - o[2, 2], o[0, 1] - o[2, 3], o[0, 2] + o[1, 3]],
                 [o[0, 1] + o[2, 3], 1.0 - o[0, 0] - o[2, 2], o[1, 2] - o[0, 3]],
                 [o[0, 2] - o[1, 3], o[1, 2] + o[0, 3], 1.0 - o[0, 0] - o[1, 1]]]
    return T


def _g6(v: float) -> str:
    """A float as the C++ ostream default (%g, six significant digits)."""
    return f"{v:g}"


def format_pose_line(tick, T: np.ndarray) -> str:
    """One pose-log line of the frame loop, from a 4x4 (float32) pose."""
    T = np.asarray(T, dtype=np.float32)
    q = quaternion_from_matrix(T[:3, :3])
    vals = [T[0, 3], T[1, 3], T[2, 3], q[0], q[1], q[2], q[3]]
    return " ".join([str(tick)] + [_g6(float(v)) for v in vals])


def write_pose_log(path: str, ticks, poses) -> None:
    with open(path, "w") as f:
        for tick, T in zip(ticks, poses):
            f.write(format_pose_line(tick, T) + "\n")


def read_trajectory(path: str) -> dict[float, np.ndarray]:
    """``stamp tx ty tz qx qy qz qw`` lines (commas and tabs allowed, '#'
    comments) to {stamp: 4x4}; lines with an all-zero quaternion or a NaN
    are skipped (evaluate_rpe.py read_trajectory)."""
    traj = {}
    with open(path) as f:
        data = f.read()
    for line in data.replace(",", " ").replace("\t", " ").split("\n"):
        if not line or line[0] == "#":
            continue
        vals = [float(v) for v in line.split(" ") if v.strip()]
        if len(vals) < 8 or vals[4:8] == [0, 0, 0, 0] or any(np.isnan(vals)):
            continue
        traj[vals[0]] = matrix_from_pose(vals[1:4], vals[4:8])
    return traj


# ------------------------------------------------------------- TUM lists


def read_file_list(path: str) -> dict[float, list[str]]:
    """``stamp d1 d2 ...`` lines to {stamp: [d1, d2, ...]} (associate.py)."""
    out = {}
    with open(path) as f:
        data = f.read()
    for line in data.replace(",", " ").replace("\t", " ").split("\n"):
        if not line or line[0] == "#":
            continue
        vals = [v.strip() for v in line.split(" ") if v.strip()]
        if len(vals) > 1:
            out[float(vals[0])] = vals[1:]
    return out


def associate(first: dict, second: dict, offset: float = 0.0, max_difference: float = 0.02) -> list:
    """Pairs (a, b) of stamps with |a - (b + offset)| < max_difference, taken
    greedily by increasing difference so that each stamp is used once, sorted
    (associate.py)."""
    cand = sorted((abs(a - (b + offset)), a, b) for a in first for b in second
                  if abs(a - (b + offset)) < max_difference)
    fa, sb = set(first), set(second)
    matches = []
    for _, a, b in cand:
        if a in fa and b in sb:
            fa.remove(a)
            sb.remove(b)
            matches.append((a, b))
    return sorted(matches)


def depth_png_to_mm(raw: np.ndarray) -> np.ndarray:
    """A TUM depth image (uint16, 1/5000 m) in millimetres as the reader
    converts it: x * (1000.0 / 5000) in double, rounded to nearest (ties to
    even), saturated to [0, 65535]."""
    v = np.rint(np.asarray(raw, dtype=np.float64) * (1000 * 1.0 / 5000))
    return np.clip(v, 0, 65535).astype(np.uint16)


@dataclass
class Frame:
    timestamp: float | int
    depth: np.ndarray            # uint16 millimetres, (h, w)
    rgb: np.ndarray              # uint8, (h, w, 3), bytes in the order the reader delivers them


class TumFrames:
    """The frames of a TUM RGB-D sequence from associated rgb / depth lists:
    line k of each list gives frame k; the number of frames is the number of
    newline characters of the rgb list, and frames 0 .. n-2 are read (the
    reader's hasMore() is currentFrame + 1 < numFrames)."""

    def __init__(self, rgb_list: str, depth_list: str, dataset_dir: str):
        from PIL import Image  # noqa: F401  (PNG/JPEG decoding)
        self.dir = dataset_dir
        with open(rgb_list) as f:
            text = f.read()
        self.num_frames = text.count("\n")
        self._rgb = [ln.split() for ln in text.splitlines() if ln.split()]
        with open(depth_list) as f:
            self._depth = [ln.split() for ln in f.read().splitlines() if ln.split()]

    def __len__(self) -> int:
        return max(self.num_frames - 1, 0)

    def __iter__(self) -> Iterator[Frame]:
        from PIL import Image
        for k in range(len(self)):
            stamp, rgb_file = self._rgb[k][0], self._rgb[k][1]
            depth_file = self._depth[k][1]
            # colour in the byte order cv::imread gives (blue, green, red),
            # which the intensity conversion weights 0.114 / 0.299 / 0.587
            rgb = np.asarray(Image.open(os.path.join(self.dir, rgb_file)).convert("RGB"))[..., ::-1].copy()
            depth = np.asarray(Image.open(os.path.join(self.dir, depth_file)))
            yield Frame(float(stamp), depth_png_to_mm(depth), rgb)


# ------------------------------------------------------------------- .klg


def read_klg(path: str, width: int, height: int, flip_colors: bool = False) -> Iterator[Frame]:
    """The frames of a ``.klg`` raw log."""
    n = width * height
    with open(path, "rb") as f:
        (count,) = struct.unpack("<i", f.read(4))
        for _ in range(count):
            ts, dsize, isize = struct.unpack("<qii", f.read(16))
            dbuf = f.read(dsize)
            ibuf = f.read(isize) if isize > 0 else b""
            if dsize == 2 * n:
                depth = np.frombuffer(dbuf, dtype="<u2").copy()
            else:
                depth = np.frombuffer(zlib.decompress(dbuf), dtype="<u2").copy()
            if isize == 3 * n:
                rgb = np.frombuffer(ibuf, dtype=np.uint8).copy()
            elif isize > 0:
                from PIL import Image
                rgb = np.asarray(Image.open(io.BytesIO(ibuf)).convert("RGB")).reshape(-1).copy()
            else:
                rgb = np.zeros(3 * n, np.uint8)
            rgb = rgb.reshape(height, width, 3)
            if flip_colors:
                rgb = rgb[..., ::-1].copy()
            yield Frame(ts, depth.reshape(height, width), rgb)


def write_klg(path: str, frames, compress_depth: bool = True, image: str = "raw") -> None:
    """Write frames (timestamp, depth uint16, rgb uint8) as a ``.klg`` log;
    image 'raw', 'jpeg' or 'none'."""
    frames = list(frames)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(frames)))
        for fr in frames:
            d = np.ascontiguousarray(fr.depth, dtype="<u2").tobytes()
            if compress_depth:
                d = zlib.compress(d)
            if image == "raw":
                im = np.ascontiguousarray(fr.rgb, dtype=np.uint8).tobytes()
            elif image == "jpeg":
                from PIL import Image
                b = io.BytesIO()
                Image.fromarray(np.ascontiguousarray(fr.rgb, dtype=np.uint8)).save(b, format="JPEG", quality=95)
                im = b.getvalue()
            else:
                im = b""
            f.write(struct.pack("<qii", int(fr.timestamp), len(d), len(im)))
            f.write(d)
            f.write(im)
