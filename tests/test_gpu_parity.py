"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer and mask outputs must be bit-identical; FP32 maps bit-identical
(including NaN positions); FP32 products summed on the device (wave/LDS trees,
FP64 across blocks) within a relative 2e-5 of the oracle's exact FP64 sums;
poses within 1 mm / 0.1 deg of the oracle's.
"""
import math

import numpy as np
import pytest

from slam_amd import synth
from slam_amd.odometry import RGBDOdometry, OdomError

from helpers import ANGLE, CUTOFF, MAXD, feed, make_pair, pose_of

pytestmark = pytest.mark.gpu

SUM_RTOL = 2e-5


def sums_close(dev, ref, rtol=SUM_RTOL):
    """Per entry, |dev - ref| <= rtol * scale where scale is the entry's
    magnitude or, for entries that cancel, the diagonal products."""
    dev, ref = np.asarray(dev, np.float64), np.asarray(ref, np.float64)
    scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    err = np.abs(dev - ref) / scale
    return float(err.max())


def mk(cam):
    return RGBDOdometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)


def orc_mk(orc, cam, fp64=False):
    return orc.Odometry(cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy, fp64=fp64)


# 202x154: odd halvings (levels 101x77, 50x38) and partial tiles at every level
CAM_ODD = synth.Camera(202, 154, 150.0, 150.0, 100.5, 76.5)


@pytest.fixture(scope="module", params=[(synth.CAM_320, 3), (synth.CAM_640, 5), (CAM_ODD, 7)],
                ids=["320x240", "640x480", "202x154"])
def pair(request, orc):
    cam, seed = request.param
    T0, T1, f0, f1 = make_pair(cam, seed)
    filt = orc.bilateral(f1.depth, MAXD)
    gpu = mk(cam)
    ref = orc_mk(orc, cam)
    feed(gpu, f0, f1, T0, filt)
    feed(ref, f0, f1, T0, filt)
    return dict(cam=cam, T0=T0, T1=T1, f0=f0, f1=f1, filt=filt, gpu=gpu, ref=ref)


def test_bilateral_bit_exact(pair, orc):
    raw = pair["f1"].depth
    out = pair["gpu"].bilateral(raw, MAXD)
    assert np.array_equal(out, pair["filt"])
    # a 3 m cut-off zeroes the far field exactly as the oracle
    assert np.array_equal(pair["gpu"].bilateral(raw, 3.0), orc.bilateral(raw, 3.0))


@pytest.mark.parametrize("maxd", [60.0, 65.1545, 70.0])
def test_bilateral_border_tiles_bit_exact(orc, maxd):
    """Border tiles at a ragged size.  Below a 65.155 m cut-off they stage
    out-of-image texels as 0xFFFF, which is more than 380 from every filtered
    value and so weighs exactly 0; above it they clip the window.  Either way
    the result is the reference's clipped window bit for bit: uniform u16
    noise, noise next to the cut-off, and a flat image at the largest value a
    65.1545 m cut-off keeps (381 from the fill)."""
    gpu = mk(CAM_ODD)
    rng = np.random.default_rng(21)
    shape = (CAM_ODD.height, CAM_ODD.width)
    for raw in (rng.integers(0, 65536, size=shape, dtype=np.uint16),
                rng.integers(64700, 65536, size=shape, dtype=np.uint16),
                np.full(shape, 65154, np.uint16)):
        ref = orc.bilateral(raw, maxd)
        assert np.array_equal(gpu.bilateral(raw, maxd), ref), maxd
    assert ref.min() == (65154 if maxd > 65.154 else 0)


def test_depth_pyramid_and_maps_bit_exact(pair):
    gpu, ref = pair["gpu"], pair["ref"]
    for l in range(3):
        assert np.array_equal(gpu.read_buffer("depth", l), ref.buf("depth_tmp", l)), l
        for name in ("vmap_curr", "nmap_curr"):
            a, b = gpu.read_buffer(name, l), ref.buf(name, l)
            assert np.array_equal(a, b, equal_nan=True), (name, l)


def test_model_maps_bit_exact(pair):
    gpu, ref = pair["gpu"], pair["ref"]
    for l in range(3):
        for name in ("vmap_g_prev", "nmap_g_prev"):
            a, b = gpu.read_buffer(name, l), ref.buf(name, l)
            assert np.array_equal(a, b, equal_nan=True), (name, l)


def test_rgb_pyramids_bit_exact(pair):
    gpu, ref = pair["gpu"], pair["ref"]
    for l in range(3):
        assert np.array_equal(gpu.read_buffer("last_depth", l), ref.buf("lastDepth", l), equal_nan=True), l
        assert np.array_equal(gpu.read_buffer("next_depth", l), ref.buf("nextDepth", l), equal_nan=True), l
        assert np.array_equal(gpu.read_buffer("last_image", l), ref.buf("lastImage", l)), l
        assert np.array_equal(gpu.read_buffer("next_image", l), ref.buf("nextImage", l)), l
        assert np.array_equal(gpu.read_buffer("last_next_image", l), ref.buf("lastNextImage", l)), l


def test_derivatives_bit_exact(pair, orc):
    gpu = pair["gpu"]
    gpu.compute_derivatives()
    for l in range(3):
        img = gpu.read_buffer("next_image", l)
        dx, dy = orc.derivative(img)
        dxy = gpu.read_buffer("dxy", l)
        assert np.array_equal(dxy[..., 0], dx), l
        assert np.array_equal(dxy[..., 1], dy), l


def _constants(T0, T1, level, cam):
    """Float constants of one iteration: current pose T1 against prior T0."""
    Rc = T1[:3, :3].astype(np.float32)
    tc = T1[:3, 3].astype(np.float32)
    Rp = T0[:3, :3].astype(np.float32)
    Rpi = np.linalg.inv(Rp.astype(np.float64)).astype(np.float32)
    tp = T0[:3, 3].astype(np.float32)
    return Rc, tc, Rpi, tp


@pytest.mark.parametrize("level", [0, 1, 2])
def test_icp_step_mask_exact_sums_close(pair, orc, level):
    gpu, ref, cam = pair["gpu"], pair["ref"], pair["cam"]
    # a pose between prior and truth, so the mask has both outcomes
    T = pair["T0"].copy()
    T[:3, 3] += 0.5 * (pair["T1"][:3, 3] - pair["T0"][:3, 3])
    Rc, tc, Rpi, tp = _constants(pair["T0"], T, level, cam)
    k = orc.level_intr(orc.intr(cam.fx, cam.fy, cam.cx, cam.cy), level)
    out_ref, m_ref = orc.icp_step(ref.buf("vmap_curr", level), ref.buf("nmap_curr", level),
                                  ref.buf("vmap_g_prev", level), ref.buf("nmap_g_prev", level), k,
                                  Rc, tc, Rpi, tp, 0.1, ANGLE)
    out, m = gpu.icp_step(level, Rc, tc, Rpi, tp)
    assert m_ref.sum() > 100
    assert np.array_equal(m, m_ref), f"{(m != m_ref).sum()} mask mismatches"
    assert out[28] == out_ref[28]
    assert sums_close(out, out_ref) < SUM_RTOL


def _warp_consts(Trel, cam, level):
    K = np.array([[cam.fx / 2 ** level, 0, cam.cx / 2 ** level], [0, cam.fy / 2 ** level, cam.cy / 2 ** level],
                  [0, 0, 1.0]])
    Rt = np.linalg.inv(Trel)
    krkinv = (K @ Rt[:3, :3] @ np.linalg.inv(K)).astype(np.float32)
    kt = (K @ Rt[:3, 3]).astype(np.float32)
    return krkinv, kt


@pytest.mark.parametrize("level", [0, 1, 2])
def test_rgb_residual_and_step(pair, orc, level):
    gpu, ref, cam = pair["gpu"], pair["ref"], pair["cam"]
    gpu.compute_derivatives()
    Trel = np.linalg.inv(pair["T0"]) @ pair["T1"]
    Trel[:3, 3] *= 0.5
    krkinv, kt = _warp_consts(Trel, cam, level)
    img = ref.buf("nextImage", level)
    dx, dy = orc.derivative(img)
    minScale = [1600.0, 576.0, 64.0][level]
    corres, cnt_ref, sig_ref = orc.rgb_residual(minScale, dx, dy, ref.buf("lastDepth", level),
                                                ref.buf("nextDepth", level), ref.buf("lastImage", level), img,
                                                krkinv, kt)
    cnt, sig, m = gpu.rgb_residual(level, krkinv, kt)
    assert cnt_ref > 50
    assert cnt == cnt_ref and sig == sig_ref
    assert np.array_equal(m, corres["valid"].astype(np.uint8))
    sigma = float(np.float32(math.sqrt(cnt)))
    k = orc.level_intr(orc.intr(cam.fx, cam.fy, cam.cx, cam.cy), level)
    out_ref = orc.rgb_step(corres, sigma, ref.buf("lastDepth", level), k, dx, dy)
    out = gpu.rgb_step(level, sigma, krkinv, kt)
    assert out[28] == out_ref[28]
    assert sums_close(out, out_ref) < SUM_RTOL


def test_so3_step(pair, orc):
    gpu, ref, cam = pair["gpu"], pair["ref"], pair["cam"]
    k2 = orc.level_intr(orc.intr(cam.fx, cam.fy, cam.cx, cam.cy), 2)
    K = np.array([[k2.fx, 0, k2.cx], [0, k2.fy, k2.cy], [0, 0, 1.0]])
    R = synth.so3_exp(np.array([0.003, -0.004, 0.002]))
    H = (K @ R @ np.linalg.inv(K)).astype(np.float32)
    Ki = np.linalg.inv(K).astype(np.float32)
    KR = (K @ R).astype(np.float32)
    out_ref, m_ref = orc.so3_step(ref.buf("lastNextImage", 2), ref.buf("nextImage", 2), H, Ki, KR)
    out, m = gpu.so3_step(H, Ki, KR)
    assert np.array_equal(m, m_ref)
    assert out[10] == out_ref[10]
    assert sums_close(out, out_ref) < SUM_RTOL


MODES = {
    "icp_rgb_so3": dict(rgbOnly=False, icpWeight=10.0, so3=True),
    "icp_only": dict(rgbOnly=False, icpWeight=100.0, so3=False),
    "rgb_only": dict(rgbOnly=True, icpWeight=10.0, so3=False),
    "icp_rgb_fast": dict(rgbOnly=False, icpWeight=10.0, so3=True, fastOdom=True),
    "icp_rgb_nopyr": dict(rgbOnly=False, icpWeight=10.0, so3=False, pyramid=False),
}


@pytest.mark.parametrize("cam", [synth.CAM_640, CAM_ODD], ids=["640x480", "202x154"])
@pytest.mark.parametrize("mode", list(MODES))
def test_track_matches_oracle(orc, mode, cam):
    T0, T1, f0, f1 = make_pair(cam, 11)
    filt = orc.bilateral(f1.depth, MAXD)
    res = {}
    for name, odo in (("gpu", mk(cam)), ("f32", orc_mk(orc, cam)), ("f64", orc_mk(orc, cam, fp64=True))):
        feed(odo, f0, f1, T0, filt)
        t, R = odo.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], **MODES[mode])
        res[name] = (pose_of(t, R), odo.stats())
    Tg = res["gpu"][0]
    for ref in ("f32", "f64"):
        dt, dr = synth.pose_error(Tg, res[ref][0])
        assert dt <= 1e-3 and dr <= 0.1, (mode, ref, dt, dr)
    if mode != "rgb_only":
        # and the tracker actually tracked
        dt, dr = synth.pose_error(Tg, T1)
        assert dt < 5e-3 and dr < 0.3, (dt, dr)
    sg, so = res["gpu"][1], res["f32"][1]
    if MODES[mode]["icpWeight"] < 100:
        assert abs(sg["lastRGBCount"] - so["lastRGBCount"]) <= 0.002 * so["lastRGBCount"] + 5
    if not MODES[mode]["rgbOnly"]:
        assert abs(sg["lastICPCount"] - so["lastICPCount"]) <= 0.002 * so["lastICPCount"] + 5


def test_full_hd_track_matches_oracle(orc):
    """1920x1080, ICP+RGB+SO3: the largest size here.  Level 0 has more tiles
    than a pass has blocks (each block loads its next tile behind the current
    gathers) and the SO3 level has more pixels than the loop's threads
    preload.  The pose agrees with the FP32 and FP64 oracles within 1 mm /
    0.1 deg, and the final ICP and photometric inlier counts with the FP32
    oracle's within 0.2 %."""
    cam = synth.Camera(1920, 1080, 1050.0, 1050.0, 959.5, 539.5)
    T0, T1, f0, f1 = make_pair(cam, 12)
    filt = orc.bilateral(f1.depth, MAXD)
    res = {}
    for name, odo in (("gpu", mk(cam)), ("f32", orc_mk(orc, cam)), ("f64", orc_mk(orc, cam, fp64=True))):
        feed(odo, f0, f1, T0, filt)
        t, R = odo.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], **MODES["icp_rgb_so3"])
        res[name] = (pose_of(t, R), odo.stats())
    Tg = res["gpu"][0]
    for ref in ("f32", "f64"):
        dt, dr = synth.pose_error(Tg, res[ref][0])
        assert dt <= 1e-3 and dr <= 0.1, (ref, dt, dr)
    dt, dr = synth.pose_error(Tg, T1)
    assert dt < 5e-3 and dr < 0.3, (dt, dr)
    sg, so = res["gpu"][1], res["f32"][1]
    assert abs(sg["lastRGBCount"] - so["lastRGBCount"]) <= 0.002 * so["lastRGBCount"] + 5
    assert abs(sg["lastICPCount"] - so["lastICPCount"]) <= 0.002 * so["lastICPCount"] + 5
    assert sg["lastICPCount"] > 1e5


def test_sequence_matches_oracle(orc):
    """Five frames of frame-to-model tracking: state carried across calls
    (lastNextImage swap, vmaps_tmp) stays in step with the oracle."""
    cam = synth.CAM_320
    scene = synth.Scene(21)
    poses = synth.spline_trajectory(6, seed=4)
    frames = [scene.render(T, cam, noise_seed=100 + i) for i, T in enumerate(poses)]
    gpu, ref = mk(cam), orc_mk(orc, cam)
    est = {"gpu": poses[0].copy(), "ref": poses[0].copy()}
    for odo in (gpu, ref):
        odo.initFirstRGB(frames[0].rgb)
    for k in range(1, len(frames)):
        filt = orc.bilateral(frames[k].depth, MAXD)
        for name, odo in (("gpu", gpu), ("ref", ref)):
            T = est[name]
            odo.initICPModel(frames[k - 1].vert4, frames[k - 1].norm4, CUTOFF, T.astype(np.float32))
            odo.initRGBModel(frames[k - 1].rgba)
            odo.initICP(filt, CUTOFF)
            odo.initRGB(frames[k].rgb)
            t, R = odo.getIncrementalTransformation(T[:3, 3], T[:3, :3], False, 10.0, True, False, True)
            est[name] = pose_of(t, R)
        dt, dr = synth.pose_error(est["gpu"], est["ref"])
        assert dt <= 1e-3 and dr <= 0.1, (k, dt, dr)
        for l in range(3):
            assert np.array_equal(gpu.read_buffer("last_next_image", l), ref.buf("lastNextImage", l))


@pytest.mark.parametrize("so3", [True, False])
def test_graph_replays_match_eager(orc, so3):
    """Six frames tracked by a handle replaying captured graphs and by one
    launching eagerly, with the frames' own init calls and a pose read-back
    (synchronisation) between the replays: every replay equals the eager
    result bit for bit (with so3 the two image parities use two graphs)."""
    cam = synth.CAM_320
    scene = synth.Scene(23)
    poses = synth.spline_trajectory(7, seed=6)
    frames = [scene.render(T, cam, noise_seed=200 + i) for i, T in enumerate(poses)]
    eager, graph = mk(cam), mk(cam)
    graph.set_graph(True)
    for odo in (eager, graph):
        odo.set_pose(poses[0][:3, 3], poses[0][:3, :3])
        odo.initFirstRGB(frames[0].rgb)
    for k in range(1, len(frames)):
        filt = orc.bilateral(frames[k].depth, MAXD)
        out = []
        for odo in (eager, graph):
            t, R = odo.get_pose()
            odo.initICPModel(frames[k - 1].vert4, frames[k - 1].norm4, CUTOFF, pose_of(t, R).astype(np.float32))
            odo.initRGBModel(frames[k - 1].rgba)
            odo.initICP(filt, CUTOFF)
            odo.initRGB(frames[k].rgb)
            odo.track_async(False, 10.0, True, False, so3)
            out.append(odo.get_pose())
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]), k


def test_degenerate_all_invalid_depth(orc):
    """No valid depth: zero information, identity update, no abort."""
    cam = synth.CAM_320
    T0, T1, f0, f1 = make_pair(cam, 2)
    gpu = mk(cam)
    zeros = np.zeros_like(f1.depth)
    gpu.initFirstRGB(f0.rgb)
    gpu.initICPModel(f0.vert4, f0.norm4, CUTOFF, T0.astype(np.float32))
    gpu.initRGBModel(f0.rgba)
    gpu.initICP(zeros, CUTOFF)
    gpu.initRGB(f1.rgb)
    t, R = gpu.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 100.0, True, False, False)
    assert np.allclose(t, T0[:3, 3]) and np.allclose(R, T0[:3, :3])
    s = gpu.stats()
    assert s["degenerate"] == 1 and s["lastICPCount"] == 0


def test_covariance_inverts_last_system(pair):
    gpu = pair["gpu"]
    T0 = pair["T0"]
    feed(gpu, pair["f0"], pair["f1"], T0, pair["filt"])
    gpu.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 10.0, True, False, True)
    A = gpu.lastA
    C = gpu.getCovariance()
    assert np.allclose(A @ C, np.eye(6), atol=1e-6)


def test_bad_arguments_raise():
    cam = synth.CAM_320
    gpu = mk(cam)
    with pytest.raises(OdomError):
        gpu.initICP(np.zeros((10, 10), np.uint16), CUTOFF)
    with pytest.raises(OdomError):
        gpu.set_step_shape(0, 0, 96, 4)


def test_call_order_is_checked():
    """initRGB* read the model vertex buffer that initICPModel / initICP(v, n)
    fill (elastic_fusion_file.cpp:461 "initICP* must be called before
    initRGB*"), and a track call needs every input its mode reads."""
    cam = synth.CAM_320
    T0, T1, f0, f1 = make_pair(cam, 4)
    gpu = mk(cam)
    with pytest.raises(OdomError) as e:
        gpu.initRGB(f1.rgb)
    assert e.value.code == -3
    with pytest.raises(OdomError) as e:
        gpu.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 100.0, True, False, False)
    assert e.value.code == -3
    gpu.initFirstRGB(f0.rgb)
    gpu.initICPModel(f0.vert4, f0.norm4, CUTOFF, T0.astype(np.float32))
    gpu.initICP(f1.depth, CUTOFF)
    # ICP-only is now complete; ICP+RGB still lacks both RGB inputs
    gpu.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 100.0, True, False, False)
    with pytest.raises(OdomError) as e:
        gpu.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 10.0, True, False, True)
    assert e.value.code == -3
    gpu.initRGBModel(f0.rgba)
    gpu.initRGB(f1.rgb)
    gpu.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 10.0, True, False, True)


def test_step_shapes_agree(pair):
    """Every launch shape of the autotuner gives the same ICP mask and sums,
    association mask and counts, and photometric sums.  The shapes cover both
    loop forms of the step kernels: one tile per block, and several (a level
    0 of 640x480 has 1200 tiles of 256 pixels for at most 256 blocks, which
    then load each next tile behind the current gathers)."""
    gpu, cam = pair["gpu"], pair["cam"]
    saved = [gpu.get_step_shape(k, 0) for k in (0, 1)]
    T = pair["T1"]
    Rc, tc, Rpi, tp = _constants(pair["T0"], T, 0, cam)
    base, mb = gpu.icp_step(0, Rc, tc, Rpi, tp)
    gpu.compute_derivatives()
    Trel = np.linalg.inv(pair["T0"]) @ T
    Trel[:3, 3] *= 0.5
    krkinv, kt = _warp_consts(Trel, cam, 0)
    cnt0, sig0, mr0 = gpu.rgb_residual(0, krkinv, kt)
    sigma = float(np.float32(math.sqrt(cnt0)))
    rgb0 = gpu.rgb_step(0, sigma, krkinv, kt)
    assert cnt0 > 50
    for t in (256, 512):
        for q in (1, 2, 3, 4):
            gpu.set_step_shape(0, 0, t, q)
            gpu.set_step_shape(1, 0, t, q)
            out, m = gpu.icp_step(0, Rc, tc, Rpi, tp)
            assert np.array_equal(m, mb)
            assert sums_close(out, base) < SUM_RTOL
            cnt, sig, mr = gpu.rgb_residual(0, krkinv, kt)
            assert (cnt, sig) == (cnt0, sig0) and np.array_equal(mr, mr0), (t, q)
            out = gpu.rgb_step(0, sigma, krkinv, kt)
            assert out[28] == rgb0[28] and sums_close(out, rgb0) < SUM_RTOL, (t, q)
    for k in (0, 1):
        gpu.set_step_shape(k, 0, *saved[k])


def test_device_memory_inputs(orc):
    torch = pytest.importorskip("torch")
    cam = synth.CAM_320
    T0, T1, f0, f1 = make_pair(cam, 8)
    filt = orc.bilateral(f1.depth, MAXD)
    host, dev = mk(cam), mk(cam)
    feed(host, f0, f1, T0, filt)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    dev.initFirstRGB(d(f0.rgb))
    dev.initICPModel(d(f0.vert4), d(f0.norm4), CUTOFF, T0.astype(np.float32))
    dev.initRGBModel(d(f0.rgba))
    dev.initICPRaw(d(f1.depth), MAXD, CUTOFF)
    dev.initRGB(d(f1.rgb))
    a = host.getIncrementalTransformation(T0[:3, 3], T0[:3, :3])
    b = dev.getIncrementalTransformation(T0[:3, 3], T0[:3, :3])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_async_device_sequence_equals_synchronous(orc):
    """Six frames enqueued from device buffers with no host wait between them
    (raw depth filtered on the device, models at the device-resident pose,
    graphs on, poses to the device log): the logged poses equal, bit for bit,
    those of the same frames fed from host buffers one synchronous call at a
    time."""
    torch = pytest.importorskip("torch")
    cam = synth.CAM_640
    scene = synth.Scene(25)
    poses = synth.spline_trajectory(7, seed=8)
    frames = [scene.render(T, cam, noise_seed=300 + i) for i, T in enumerate(poses)]
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    dev = [{k: d(getattr(f, k)) for k in ("depth", "rgb", "rgba", "vert4", "norm4")} for f in frames]
    torch.cuda.synchronize()
    sync, asy = mk(cam), mk(cam)
    asy.set_graph(True)
    log = torch.zeros((len(frames), 12), dtype=torch.float32, device="cuda")
    asy.set_pose_log(log.data_ptr(), len(frames))
    for odo in (sync, asy):
        odo.set_pose(poses[0][:3, 3], poses[0][:3, :3])
    sync.initFirstRGB(frames[0].rgb)
    asy.initFirstRGB(dev[0]["rgb"])
    expected = []
    for k in range(1, len(frames)):
        sync.initICPModelCurrent(frames[k - 1].vert4, frames[k - 1].norm4, CUTOFF)
        sync.initRGBModel(frames[k - 1].rgba)
        sync.initICPRaw(frames[k].depth, MAXD, CUTOFF)
        sync.initRGB(frames[k].rgb)
        sync.track_async(False, 10.0, True, False, True)
        expected.append(sync.get_pose())
        asy.initICPModelCurrent(dev[k - 1]["vert4"], dev[k - 1]["norm4"], CUTOFF)
        asy.initRGBModel(dev[k - 1]["rgba"])
        asy.initICPRaw(dev[k]["depth"], MAXD, CUTOFF)
        asy.initRGB(dev[k]["rgb"])
        asy.track_async(False, 10.0, True, False, True)
    asy.synchronize()
    got = log.cpu().numpy()
    for k, (t, R) in enumerate(expected):
        Rt = got[k].reshape(3, 4)
        assert np.array_equal(Rt[:, 3], np.asarray(t, np.float32).reshape(3)), k
        assert np.array_equal(Rt[:, :3], np.asarray(R, np.float32).reshape(3, 3)), k
