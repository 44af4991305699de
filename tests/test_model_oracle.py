"""The surfel fusion oracle (oracle/model_oracle.c) pinned by what the
passes must give on scenes whose answer is known: a fronto-parallel wall
initialises one surfel per pixel with the closed-form position, normal,
radius and confidence; the index map puts each surfel at the pixel it
projects to, nearest first and the lower index on ties; fusing the same
frame again associates every selected pixel with its own surfel and adds the
confidences; a wall 10 cm away associates with nothing and comes back as new
points; the cleaning drops stale unconfident surfels and keeps confident or
long-unseen ones."""
import numpy as np
import pytest

W, H = 32, 24
FX = FY = 30.0
CX, CY = 15.5, 11.5
MAXD = 5.0          # depth_metric cutoff (m)
MAX_DEPTH = 20.0    # maxDepthProcessed


def orc():
    from oracle import oracle
    return oracle


def wall(z_m):
    return np.full((H, W), int(round(z_m * 1000)), np.uint16)


def rgb_img(seed=0):
    return np.random.default_rng(seed).integers(0, 256, (H, W, 3), dtype=np.uint8)


def model(max_surfels=4096):
    return orc().SurfelModel(W, H, CX, CY, FX, FY, max_surfels)


def test_encode_color_is_exact_for_bytes():
    b = np.arange(256, dtype=np.float32)
    assert np.array_equal(np.round((b / np.float32(255.0)) * np.float32(255.0)), b)


def test_depth_metric_window():
    d = np.array([0, 299, 300, 1234, 5000, 5001, 65535], np.uint16)
    assert orc().depth_metric(d, 5.0).tolist() == pytest.approx([0, 0, 0.3, 1.234, 5.0, 0, 0])


def test_initialise_a_wall():
    m = model()
    rgb = rgb_img(1)
    z = 2.0
    n = m.initialise(rgb, wall(z), wall(z), MAXD, 7, MAX_DEPTH)
    assert n == W * H
    M = m.model()
    # column-major draw order: surfel k is pixel (k // H, k % H)
    i, j = np.arange(n) // H, np.arange(n) % H
    x, y = i + 0.5, j + 0.5
    assert np.allclose(M[:, 0], (x - CX) * z / FX, atol=1e-5)
    assert np.allclose(M[:, 1], (y - CY) * z / FY, atol=1e-5)
    assert np.all(M[:, 2] == np.float32(z))
    conf = np.exp(-((np.hypot(x - CX, y - CY) / 400.0) ** 2) / 0.72)
    assert np.allclose(M[:, 3], conf, rtol=1e-6)
    col = (rgb[j, i, 0].astype(np.int64) << 16) | (rgb[j, i, 1].astype(np.int64) << 8) | rgb[j, i, 2]
    assert np.array_equal(M[:, 4], col.astype(np.float32))
    assert np.all(M[:, 5] == 0) and np.all(M[:, 6] == 1) and np.all(M[:, 7] == 7)
    # the wall's normal (away from the camera, as the central differences give it)
    assert np.allclose(M[:, 8:11], [0, 0, 1], atol=1e-6)
    assert np.allclose(M[:, 11], z / FX * np.sqrt(2), rtol=1e-6)
    # pixels without depth make no surfel
    d = wall(z)
    d[5, 7] = 0
    d[10, 3] = 100            # below the 0.3 m window
    assert model().initialise(rgb, d, wall(z), MAXD, 7, MAX_DEPTH) == W * H - 2


def test_full_buffer_keeps_the_first_surfels():
    """Transform feedback into a full buffer drops what does not fit: a model
    of capacity 100 holds the first 100 surfels of the unbounded one."""
    rgb = rgb_img(4)
    d = wall(2.0)              # every pixel makes a surfel; the random colours tell them apart
    full, small = model(4096), model(100)
    n = full.initialise(rgb, d, d, MAXD, 3, MAX_DEPTH)
    assert n == W * H and small.initialise(rgb, d, d, MAXD, 3, MAX_DEPTH) == 100
    assert np.array_equal(small.model(), full.model()[:100])


def one(x, y, z, k_conf=5.0, last=3.0, init=1.0, n=(0.0, 0.0, 1.0), r=0.01, col=0.0):
    return [x, y, z, k_conf, col, 0.0, init, last, n[0], n[1], n[2], r]


def test_index_map_nearest_and_ties():
    m = model()
    # pixel (i, j) = (20, 7): ray through its centre
    u, v = (20 + 0.5 - CX) / FX, (7 + 0.5 - CY) / FY
    S = [one(9, 9, -1.0),                       # surfel 0 (behind the camera)
         one(u * 2.0, v * 2.0, 2.0, col=1.0),   # 1: at 2 m
         one(u * 1.5, v * 1.5, 1.5, col=2.0),   # 2: nearer
         one(u * 1.5, v * 1.5, 1.5, col=3.0),   # 3: tie with 2
         one(u * 3.0, v * 3.0, 3.0, last=-500)]  # 4: too old for the time window
    m.set(np.array(S, np.float32))
    m.predictIndices(np.eye(4), 10, MAX_DEPTH, 200)
    idx, vc, ct, nr = m.index_map()
    assert idx[7, 20] == 2 and np.count_nonzero(idx) == 1
    assert ct[7, 20, 0] == 2.0
    assert np.allclose(vc[7, 20], [u * 1.5, v * 1.5, 1.5, 5.0])
    # surfel 0 writes index 0, the value that means "none" (a quirk kept)
    m.set(np.array([one(u * 2.0, v * 2.0, 2.0)], np.float32))
    m.predictIndices(np.eye(4), 10, MAX_DEPTH, 200)
    idx, vc, _, _ = m.index_map()
    assert not idx.any() and vc[7, 20, 2] == np.float32(2.0)


def test_fusing_the_same_frame_updates_every_selected_pixel_in_place():
    m = model()
    rgb = rgb_img(2)
    z = 2.0
    n = m.initialise(rgb, wall(z), wall(z), MAXD, 1, MAX_DEPTH)
    before = m.model()
    T = np.eye(4)
    m.predictIndices(T, 2, MAX_DEPTH, 200)
    m.fuse(T, 2, rgb, wall(z), wall(z), MAXD, MAX_DEPTH, 1.0)
    after = m.model()
    fresh = m.fresh()
    i, j = np.arange(n) // H, np.arange(n) % H
    # selected: both coordinates even (time 2) with four neighbours of depth;
    # sampling clamps to the edge, so the border's neighbours are itself
    sel = (i % 2 == 0) & (j % 2 == 0)
    assert fresh.shape[0] == sel.sum() and np.all(fresh[:, 7] == -1)       # all matched
    # each matched its own surfel (the nearest ray): confidence doubled,
    # position and colour kept.  Surfel 0 is the exception: the index map
    # writes its index 0, which reads as "none", so pixel (0, 0) matches the
    # surfel of its 3x3 window nearest to its ray.
    own = sel.copy()
    own[0] = False
    assert np.allclose(after[own, 3], 2 * before[own, 3], rtol=1e-6)
    assert np.allclose(after[own, :3], before[own, :3], atol=1e-6)
    assert np.array_equal(after[own, 4], before[own, 4])
    assert np.all(after[own, 7] == 2)
    ray = np.array([(0.5 - CX) / FX, (0.5 - CY) / FY, 1.0])
    cand = [1, H, H + 1]                       # pixels (0, 1), (1, 0), (1, 1)
    dist = [np.linalg.norm(np.cross(ray, before[k, :3])) / np.linalg.norm(ray) for k in cand]
    k1 = cand[int(np.argmin(dist))]
    assert np.allclose(after[k1, 3], before[k1, 3] + before[0, 3], rtol=1e-6) and after[k1, 7] == 2
    untouched = ~own
    untouched[k1] = False
    assert np.array_equal(after[untouched], before[untouched])
    # cleaning drops the matched copies (w = -1) and keeps the model
    m.predictIndices(T, 2, MAX_DEPTH, 200)
    assert m.clean(T, 2, 10.0, 200) == n


def test_a_new_wall_comes_back_as_new_points():
    m = model()
    rgb = rgb_img(3)
    n = m.initialise(rgb, wall(2.0), wall(2.0), MAXD, 1, MAX_DEPTH)
    T = np.eye(4)
    m.predictIndices(T, 2, MAX_DEPTH, 200)
    m.fuse(T, 2, rgb, wall(2.1), wall(2.1), MAXD, MAX_DEPTH, 1.0)
    fresh = m.fresh()
    assert fresh.shape[0] > 0 and np.all(fresh[:, 7] == -2)
    assert np.all(fresh[:, 2] == np.float32(2.1)) and np.all(fresh[:, 6] == 2)
    # draw order: column by column
    cols = np.round(fresh[:, 0] * FX / 2.1 + CX - 0.5).astype(int)
    rows = np.round(fresh[:, 1] * FY / 2.1 + CY - 0.5).astype(int)
    order = cols * H + rows
    assert np.all(np.diff(order) > 0)
    m.predictIndices(T, 2, MAX_DEPTH, 200)
    count = m.clean(T, 2, 10.0, 200)
    assert count == n + fresh.shape[0]
    assert np.all(m.model()[n:, 7] == 2)        # new points stamped with the time


def test_cleaning_rules():
    m = model()
    u, v = (10.5 - CX) / FX, (5.5 - CY) / FY
    S = np.array([one(u, v, 1.0, k_conf=1.0, last=1.0),     # stale and unconfident: dropped
                  one(u, v, 1.0, k_conf=20.0, last=1.0),    # confident: kept
                  one(u, v, 1.0, k_conf=1.0, last=25.0),    # seen recently: kept
                  one(u, v, 1.0, k_conf=1.0, last=-1.0),    # a matched copy: dropped
                  ], np.float32)
    m.set(S)
    T = np.eye(4)
    m.predictIndices(T, 30, MAX_DEPTH, 200)
    assert m.clean(T, 30, 10.0, 200) == 2
    kept = m.model()
    assert kept[0, 3] == 20.0 and kept[1, 7] == 25.0
    # beyond timeDelta the stale surfel is kept (inactive, not cleaned)
    m.set(S[:1])
    m.predictIndices(T, 300, MAX_DEPTH, 200)
    assert m.clean(T, 300, 10.0, 200) == 1
