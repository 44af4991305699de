"""Parity of the HIP surfel fusion (through the C ABI) with the CPU oracle,
bit for bit, over short sequences: the first frame's initialisation, then per
frame the index map, the fusion (updated model and new points), the index map
again and the cleaning (kept surfels, in order) -- at 320x240, 640x480 and an
odd 202x154, with the frames' own depth noise and bilateral filtering.  The
fused model then feeds the splat prediction on the device."""
import numpy as np
import pytest

from slam_amd import splat, surfels, synth

from helpers import MAXD

pytestmark = pytest.mark.gpu

CAM_ODD = synth.Camera(202, 154, 150.0, 150.0, 100.5, 76.5)
DEPTH_CUTOFF = 5.0      # depth_metric window (m)
MAX_DEPTH = 20.0        # maxDepthProcessed
CONF = 10.0
DELTA = 200


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def sequence(cam, seed, n):
    """Frames of one room along a smooth path, poses relative to the first
    (the first frame's pose is the identity, as the model's frame)."""
    scene = synth.Scene(seed)
    poses = synth.spline_trajectory(n, seed=seed, max_step_m=0.02, max_step_deg=1.0)
    T0i = np.linalg.inv(poses[0])
    out = []
    for k, T in enumerate(poses):
        f = scene.render(T, cam, noise_seed=seed * 100 + k)
        out.append(((T0i @ T).astype(np.float32), f))
    return out


def compare_index_maps(gpu, ref, what):
    for name, g, e in zip(("index", "vert_conf", "color_time", "norm_rad"), gpu.index_map(), ref.index_map()):
        assert same_bits(g, e), (what, name, int(np.sum(g != e)))


@pytest.mark.parametrize("cam,seed,frames", [(synth.CAM_320, 61, 6), (synth.CAM_640, 62, 3), (CAM_ODD, 63, 5)],
                         ids=["320x240", "640x480", "202x154"])
def test_fusion_sequence_bit_exact(orc, cam, seed, frames):
    seq = sequence(cam, seed, frames)
    args = (cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    cap = cam.width * cam.height * (frames + 1)
    gpu = surfels.SurfelModel(*args, cap)
    ref = orc.SurfelModel(*args, cap)
    T, f = seq[0]
    filt = orc.bilateral(f.depth, MAXD)
    ng = gpu.initialise(f.rgb, f.depth, filt, DEPTH_CUTOFF, 1, MAX_DEPTH)
    nr = ref.initialise(f.rgb, f.depth, filt, DEPTH_CUTOFF, 1, MAX_DEPTH)
    assert ng == nr and ng > 0.5 * cam.width * cam.height
    assert same_bits(gpu.model(), ref.model())
    for k in range(1, frames):
        tick = k + 1
        T, f = seq[k]
        filt = orc.bilateral(f.depth, MAXD)
        weighting = 1.0 - 0.1 * k
        for m in (gpu, ref):
            m.predictIndices(T, tick, MAX_DEPTH, DELTA)
        compare_index_maps(gpu, ref, ("first", k))
        for m in (gpu, ref):
            m.fuse(T, tick, f.rgb, f.depth, filt, DEPTH_CUTOFF, MAX_DEPTH, weighting)
        assert same_bits(gpu.model(), ref.model()), ("fused model", k)
        fg, fr = gpu.fresh(), ref.fresh()
        assert fg.shape == fr.shape and same_bits(fg, fr), ("new points", k)
        assert (fr[:, 7] == -1).sum() > 0 and (fr[:, 7] == -2).sum() > 0
        for m in (gpu, ref):
            m.predictIndices(T, tick, MAX_DEPTH, DELTA)
        compare_index_maps(gpu, ref, ("second", k))
        cg = gpu.clean(T, tick, CONF, DELTA)
        cr = ref.clean(T, tick, CONF, DELTA)
        assert cg == cr, (k, cg, cr)
        assert same_bits(gpu.model(), ref.model()), ("cleaned model", k)


@pytest.mark.parametrize("frac", [0.35, 0.979], ids=["initialise-overflows", "clean-overflows"])
def test_full_model_keeps_the_first_surfels_in_order(orc, frac):
    """A model buffer smaller than the sequence needs: initialise and clean
    keep the first max_surfels surfels in draw order and drop the rest, as
    the oracle does (transform feedback into a full buffer)."""
    cam = synth.CAM_320
    seq = sequence(cam, 64, 3)
    args = (cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    cap = int(frac * cam.width * cam.height)
    gpu = surfels.SurfelModel(*args, cap)
    ref = orc.SurfelModel(*args, cap)
    T, f = seq[0]
    filt = orc.bilateral(f.depth, MAXD)
    ng = gpu.initialise(f.rgb, f.depth, filt, DEPTH_CUTOFF, 1, MAX_DEPTH)
    nr = ref.initialise(f.rgb, f.depth, filt, DEPTH_CUTOFF, 1, MAX_DEPTH)
    assert ng == nr and ng <= cap and same_bits(gpu.model(), ref.model())
    full = ng == cap
    for k in range(1, len(seq)):
        T, f = seq[k]
        filt = orc.bilateral(f.depth, MAXD)
        for m in (gpu, ref):
            m.predictIndices(T, k + 1, MAX_DEPTH, DELTA)
            m.fuse(T, k + 1, f.rgb, f.depth, filt, DEPTH_CUTOFF, MAX_DEPTH, 1.0)
            m.predictIndices(T, k + 1, MAX_DEPTH, DELTA)
        assert same_bits(gpu.model(), ref.model()), ("fused model", k)
        cg = gpu.clean(T, k + 1, CONF, DELTA)
        cr = ref.clean(T, k + 1, CONF, DELTA)
        assert cg == cr and cg <= cap, (k, cg, cr)
        assert same_bits(gpu.model(), ref.model()), ("cleaned model", k)
        full = full or cg == cap
    assert full                      # the capacity was reached


def test_cleaning_drops_stale_surfels_on_the_device(orc):
    """Thirty frames past the last update, unconfident surfels go and
    confident ones stay, as the oracle decides."""
    cam = synth.CAM_320
    seq = sequence(cam, 64, 1)
    T, f = seq[0]
    filt = orc.bilateral(f.depth, MAXD)
    args = (cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    gpu = surfels.SurfelModel(*args, 200000)
    ref = orc.SurfelModel(*args, 200000)
    for m in (gpu, ref):
        m.initialise(f.rgb, f.depth, filt, DEPTH_CUTOFF, 1, MAX_DEPTH)
    M = gpu.model()
    M[::3, 3] = 50.0            # every third surfel confident
    for m in (gpu, ref):
        m.set_model(M)
        m.predictIndices(T, 31, MAX_DEPTH, DELTA)
    n = gpu.clean(T, 31, CONF, DELTA)
    assert n == ref.clean(T, 31, CONF, DELTA) == (M[:, 3] > CONF).sum()
    assert same_bits(gpu.model(), ref.model())


def test_fused_model_feeds_the_splat_prediction(orc):
    import torch
    cam = synth.CAM_320
    seq = sequence(cam, 65, 3)
    args = (cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    gpu = surfels.SurfelModel(*args, 400000)
    ref = orc.SurfelModel(*args, 400000)
    for k, (T, f) in enumerate(seq):
        filt = orc.bilateral(f.depth, MAXD)
        for m in (gpu, ref):
            if k == 0:
                m.initialise(f.rgb, f.depth, filt, DEPTH_CUTOFF, 1, MAX_DEPTH)
            else:
                m.predictIndices(T, k + 1, MAX_DEPTH, DELTA)
                m.fuse(T, k + 1, f.rgb, f.depth, filt, DEPTH_CUTOFF, MAX_DEPTH, 1.0)
                m.predictIndices(T, k + 1, MAX_DEPTH, DELTA)
                m.clean(T, k + 1, CONF, DELTA)
    # the device model, as the splat's input without a copy
    ptr, n = gpu.device_model()
    model_host = gpu.model()
    assert same_bits(model_host, ref.model())
    dev = torch.from_numpy(model_host).cuda()      # same bytes, to compare the pointer path with
    im = splat.IndexMap(*args)
    T = seq[-1][0]
    im.combinedPredict(T, dev, MAX_DEPTH, 0.0, 3, 3, DELTA)
    exp = orc.splat_predict(*args, T, model_host, MAX_DEPTH, 0.0, 3, 3, DELTA)
    assert same_bits(im.read(splat.VERTEX), exp[1])
    assert (exp[1][..., 2] > 0).mean() > 0.5
    assert ptr != 0 and n == model_host.shape[0]


def test_bad_arguments_raise():
    from slam_amd.odometry import OdomError
    cam = synth.CAM_320
    args = (cam.width, cam.height, cam.cx, cam.cy, cam.fx, cam.fy)
    m = surfels.SurfelModel(*args, 1000)
    with pytest.raises(OdomError):                       # more surfels than the model holds
        m.set_model(np.zeros((1001, 12), np.float32))
    with pytest.raises(OdomError):                       # frame of another size
        m.initialise(np.zeros((10, 10, 3), np.uint8), np.zeros((10, 10), np.uint16),
                     np.zeros((10, 10), np.uint16), DEPTH_CUTOFF, 1, MAX_DEPTH)
    with pytest.raises(OdomError):                       # max_surfels + pixels beyond 2^23
        surfels.SurfelModel(*args, 1 << 23)
    # an empty model: the index map is empty and cleaning keeps nothing
    m.set_model(np.zeros((0, 12), np.float32))
    T = np.eye(4, dtype=np.float32)
    m.predictIndices(T, 1, MAX_DEPTH, DELTA)
    assert not m.index_map()[0].any()
    assert m.clean(T, 1, CONF, DELTA) == 0
