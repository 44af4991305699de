"""Parity of the HIP splat prediction and fill-in (through the C ABI) with
the CPU oracle: every output bit-identical (float maps including their NaN
positions), on models with heavy overlap, exact duplicates (ties), odd image
sizes, the active and inactive time windows, host and device models, an
empty model, and the prediction -> fill-in -> odometry chain on the device."""
import numpy as np
import pytest

from slam_amd import splat, synth
from slam_amd.odometry import RGBDOdometry

from helpers import CUTOFF, MAXD, pose_of

pytestmark = pytest.mark.gpu

CAM_ODD = synth.Camera(202, 154, 150.0, 150.0, 100.5, 76.5)


def cam_args(cam):
    return cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy


def model_of(cam, seed, frames=3, step=1):
    """Surfels from a few views of one scene (overlapping coverage), with
    jittered confidence and a spread of last-seen times."""
    scene = synth.Scene(seed)
    poses = synth.spline_trajectory(frames, seed=seed, max_step_m=0.05, max_step_deg=3.0)
    parts = []
    for k, T in enumerate(poses):
        f = scene.render(T, cam, noise_seed=seed * 10 + k)
        parts.append(synth.surfels_from_frame(f, cam, step=step, conf=5.0 + k, init_time=k, last_time=10 * k,
                                              conf_jitter=4.0, seed=seed + k))
    return np.concatenate(parts), poses, scene


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def check_prediction(im, cam, pose, model, **kw):
    ref = __import__("oracle.oracle", fromlist=["oracle"])
    args = dict(depthCutoff=CUTOFF, confThreshold=6.0, time=20, maxTime=20, timeDelta=15)
    args.update(kw)
    exp = ref.splat_predict(*cam_args(cam), pose, model, args["depthCutoff"], args["confThreshold"], args["time"],
                            args["maxTime"], args["timeDelta"])
    got = [im.read(splat.IMAGE), im.read(splat.VERTEX), im.read(splat.NORMAL), im.read(splat.TIME)]
    for name, g, e in zip(("image", "vertex", "normal", "time"), got, exp):
        assert same_bits(g, e), (name, int(np.sum(g != e)))
    return exp


def test_prediction_bit_exact_on_hard_geometry():
    """Surfels placed at random in the view: distances 0.15-4 m, normals
    uniform on the sphere (so some are seen edge-on or from behind), radii
    log-uniform 1 mm-15 cm (sprites up to the 256-pixel point size), with
    duplicates.  The raster's candidate bound and its fall-back to the whole
    sprite square (grazing views, radii above 2 % of the distance, disks that
    reach close to the camera plane) must keep every passing fragment: the
    prediction equals the oracle's bit for bit."""
    cam = CAM_ODD
    rng = np.random.default_rng(47)
    n = 6000
    z = rng.uniform(0.15, 4.0, n)
    x = (rng.uniform(0, cam.width, n) - cam.cx) / cam.fx * z
    y = (rng.uniform(0, cam.height, n) - cam.cy) / cam.fy * z
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    # a fifth of the normals nearly perpendicular to the view ray (edge-on disks)
    ray = np.stack([x, y, z], 1) / np.linalg.norm(np.stack([x, y, z], 1), axis=1, keepdims=True)
    k = n // 5
    side = np.cross(ray[:k], nrm[:k])
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    nrm[:k] = side + rng.uniform(-0.05, 0.05, (k, 1)) * ray[:k]
    nrm[:k] /= np.linalg.norm(nrm[:k], axis=1, keepdims=True)
    model = np.zeros((n, 12), np.float32)
    model[:, 0], model[:, 1], model[:, 2] = x, y, z
    model[:, 3] = rng.uniform(7.0, 12.0, n)
    model[:, 4] = rng.integers(0, 1 << 24, n).astype(np.float32)
    model[:, 6] = rng.integers(0, 10, n)
    model[:, 7] = rng.integers(10, 20, n)
    model[:, 8:11] = nrm
    model[:, 11] = np.exp(rng.uniform(np.log(0.001), np.log(0.15), n))
    model = np.concatenate([model, model[:500]])          # exact duplicates (ties)
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    pose = np.eye(4, dtype=np.float32)
    im.combinedPredict(pose, model, CUTOFF, 6.0, 20, 20, 15)
    exp = check_prediction(im, cam, pose, model)
    assert (exp[1][..., 2] > 0).mean() > 0.5


@pytest.mark.parametrize("cam,seed", [(synth.CAM_320, 41), (synth.CAM_640, 42), (CAM_ODD, 43)],
                         ids=["320x240", "640x480", "202x154"])
def test_prediction_bit_exact(cam, seed):
    model, poses, _ = model_of(cam, seed)
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    pose = poses[1].astype(np.float32)
    im.combinedPredict(pose, model, CUTOFF, 6.0, 20, 20, 15)
    exp = check_prediction(im, cam, pose, model)
    assert (exp[1][..., 2] > 0).mean() > 0.5     # the model covers the view


def test_duplicates_resolve_to_the_first_drawn():
    """Every surfel three times, the copies with other colours and times:
    each pixel shows the first copy, as GL_LESS in draw order does."""
    cam = synth.CAM_320
    model, poses, _ = model_of(cam, 44, frames=1)
    n = model.shape[0]
    b, c = model.copy(), model.copy()
    b[:, 4], b[:, 6] = 0x00FF00, 1000
    c[:, 4], c[:, 6] = 0x0000FF, 2000
    dup = np.concatenate([model, b, c])
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=dup.shape[0])
    pose = poses[0].astype(np.float32)
    im.combinedPredict(pose, dup, CUTOFF, 0.0, 0, 0, 100)
    exp = check_prediction(im, cam, pose, dup, confThreshold=0.0, time=0, maxTime=0, timeDelta=100)
    tm = exp[3]
    assert n > 0 and not np.any(tm >= 1000)


@pytest.mark.parametrize("window", ["active", "inactive"])
def test_time_windows(window):
    cam = synth.CAM_320
    model, poses, _ = model_of(cam, 45)
    tick, delta = 25, 12
    kw = dict(time=tick, maxTime=tick, timeDelta=delta) if window == "active" else \
        dict(time=0, maxTime=tick - delta, timeDelta=delta)
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    pose = poses[2].astype(np.float32)
    im.combinedPredict(pose, model, CUTOFF, 6.0, kw["time"], kw["maxTime"], kw["timeDelta"])
    exp = check_prediction(im, cam, pose, model, **kw)
    assert exp[3][exp[1][..., 2] > 0].size > 0


def test_device_model_equals_host_model_and_empty_model_clears():
    import torch
    cam = synth.CAM_320
    model, poses, _ = model_of(cam, 46)
    pose = poses[1].astype(np.float32)
    host = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    dev = splat.IndexMap(*cam_args(cam))
    host.combinedPredict(pose, model, CUTOFF, 6.0, 20, 20, 15)
    dev.combinedPredict(pose, torch.from_numpy(model).cuda(), CUTOFF, 6.0, 20, 20, 15)
    for w in (splat.IMAGE, splat.VERTEX, splat.NORMAL, splat.TIME):
        assert same_bits(host.read(w), dev.read(w))
    # a second prediction of another pose on the same handle: the key buffer
    # was cleared by the first
    pose2 = poses[2].astype(np.float32)
    dev.combinedPredict(pose2, torch.from_numpy(model).cuda(), CUTOFF, 6.0, 20, 20, 15)
    check_prediction(dev, cam, pose2, model)
    dev.combinedPredict(pose2, np.zeros((0, 12), np.float32), CUTOFF, 6.0, 20, 20, 15)
    for w in (splat.IMAGE, splat.VERTEX, splat.NORMAL, splat.TIME):
        assert not dev.read(w).any()


@pytest.mark.parametrize("cam", [synth.CAM_320, CAM_ODD], ids=["320x240", "202x154"])
@pytest.mark.parametrize("passthrough", [(False, False), (True, False), (False, True)])
def test_fill_in_bit_exact(cam, passthrough):
    ref = __import__("oracle.oracle", fromlist=["oracle"])
    model, poses, scene = model_of(cam, 47, step=2)
    T = poses[2]
    f = scene.render(T, cam, noise_seed=5)
    filt = ref.bilateral(f.depth, MAXD)
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    im.combinedPredict(T.astype(np.float32), model, CUTOFF, 6.0, 20, 20, 15)
    im.fillIn(filt, f.rgb, *passthrough)
    pred = [im.read(splat.IMAGE), im.read(splat.VERTEX), im.read(splat.NORMAL)]
    exp = ref.splat_fill_in(*cam_args(cam), *pred, filt, f.rgb, *passthrough)
    got = [im.read(splat.FILL_IMAGE), im.read(splat.FILL_VERTEX), im.read(splat.FILL_NORMAL)]
    for name, g, e in zip(("image", "vertex", "normal"), got, exp):
        assert same_bits(g, e), (name, int(np.sum(g != e)))


def test_prediction_feeds_the_odometry_on_the_device():
    """The frame loop's predict() then frame-to-model tracking: the fill-in
    maps handed over as device images track exactly as the same maps copied
    to the host and passed back."""
    ref = __import__("oracle.oracle", fromlist=["oracle"])
    cam = synth.CAM_320
    model, poses, scene = model_of(cam, 48)
    T0 = poses[2]
    T1 = synth.perturb(T0, np.random.default_rng(3), 0.02, 1.0)
    f0 = scene.render(T0, cam, noise_seed=7)
    f1 = scene.render(T1, cam, noise_seed=8)
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    im.combinedPredict(T0.astype(np.float32), model, CUTOFF, 6.0, 20, 20, 15)
    im.fillIn(ref.bilateral(f0.depth, MAXD), f0.rgb)
    filt1 = ref.bilateral(f1.depth, MAXD)
    out = []
    for device in (True, False):
        odo = RGBDOdometry(*cam_args(cam)[:2], cam.cx, cam.cy, cam.fx, cam.fy)
        odo.initFirstRGB(f0.rgb)
        if device:
            v, n, c = im.device(splat.FILL_VERTEX), im.device(splat.FILL_NORMAL), im.device(splat.FILL_IMAGE)
        else:
            v, n, c = im.read(splat.FILL_VERTEX), im.read(splat.FILL_NORMAL), im.read(splat.FILL_IMAGE)
        odo.initICPModel(v, n, CUTOFF, T0.astype(np.float32))
        odo.initRGBModel(c)
        odo.initICP(filt1, CUTOFF)
        odo.initRGB(f1.rgb)
        t, R = odo.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 10.0, True, False, True)
        out.append(pose_of(t, R))
    assert np.array_equal(out[0], out[1])
    dt, dr = synth.pose_error(out[0], T1)
    assert dt < 1e-2 and dr < 0.5, (dt, dr)
    # the prediction enqueued on the odometry's own stream: ordered by the
    # stream alone, no host wait between the prediction and the init calls
    odo = RGBDOdometry(*cam_args(cam)[:2], cam.cx, cam.cy, cam.fx, cam.fy)
    im2 = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    im2.set_stream(odo.stream)
    odo.initFirstRGB(f0.rgb)
    im2.combinedPredict(T0.astype(np.float32), model, CUTOFF, 6.0, 20, 20, 15)
    im2.fillIn(ref.bilateral(f0.depth, MAXD), f0.rgb)
    odo.initICPModel(im2.device(splat.FILL_VERTEX), im2.device(splat.FILL_NORMAL), CUTOFF, T0.astype(np.float32))
    odo.initRGBModel(im2.device(splat.FILL_IMAGE))
    odo.initICP(filt1, CUTOFF)
    odo.initRGB(f1.rgb)
    t, R = odo.getIncrementalTransformation(T0[:3, 3], T0[:3, :3], False, 10.0, True, False, True)
    assert np.array_equal(pose_of(t, R), out[0])
    im2.set_stream(None)


def test_bad_arguments_raise():
    from slam_amd.odometry import OdomError
    cam = synth.CAM_320
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=10)
    pose = np.eye(4, dtype=np.float32)
    with pytest.raises(OdomError):                       # more surfels than the host staging holds
        im.combinedPredict(pose, np.zeros((11, 12), np.float32), CUTOFF, 0.0, 0, 0, 100)
    with pytest.raises(ValueError):
        im.combinedPredict(pose, np.zeros((5, 11), np.float32), CUTOFF, 0.0, 0, 0, 100)
    with pytest.raises(OdomError):                       # frame of another size
        im.fillIn(np.zeros((10, 10), np.uint16), np.zeros((10, 10, 3), np.uint8))
    with pytest.raises(OdomError):
        splat.IndexMap(0, 10, 1.0, 1.0, 1.0, 1.0)
    # an empty model is a valid prediction (all outputs cleared)
    im.combinedPredict(pose, np.zeros((0, 12), np.float32), CUTOFF, 0.0, 0, 0, 100)
    assert not im.read(splat.VERTEX).any()


def test_large_sprites_shared_by_the_wave(orc):
    """Sprites above the per-lane limit (here up to ~130 px wide) are worked
    through by the whole wave: same outputs as the oracle, mixed with small
    ones in the same waves and overlapping each other."""
    cam = synth.CAM_320
    rng = np.random.default_rng(5)
    model = []
    for k in range(300):
        z = rng.uniform(0.8, 4.0)
        big = k % 7 == 0
        r = rng.uniform(0.1, 0.3) if big else rng.uniform(0.002, 0.02)
        n = rng.normal(size=3)
        n[2] = -abs(n[2]) - 0.5
        n /= np.linalg.norm(n)
        model.append([rng.uniform(-0.6, 0.6) * z, rng.uniform(-0.45, 0.45) * z, z, 20.0,
                      float(rng.integers(0, 1 << 24)), 0.0, float(k), 5.0, n[0], n[1], n[2], r])
    model = np.array(model, np.float32)
    im = splat.IndexMap(*cam_args(cam), max_host_surfels=model.shape[0])
    pose = np.eye(4, dtype=np.float32)
    im.combinedPredict(pose, model, CUTOFF, 0.0, 5, 5, 100)
    exp = orc.splat_predict(*cam_args(cam), pose, model, CUTOFF, 0.0, 5, 5, 100)
    for name, w, e in zip(("image", "vertex", "normal", "time"), (splat.IMAGE, splat.VERTEX, splat.NORMAL, splat.TIME),
                          exp):
        g = im.read(w)
        assert same_bits(g, e), (name, int(np.sum(g != e)))
    assert (exp[1][..., 2] > 0).mean() > 0.3
