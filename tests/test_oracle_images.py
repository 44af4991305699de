"""Pins of the oracle's image operations (a1-a13) on the CPU.

Every check comes from outside the oracle's own formula: constant images that a
normalised filter must leave unchanged, closed forms on planes and ramps,
textbook routines (math.exp, a 5x5 binomial correlation) in the interior, and
special values (invalid depths, NaN propagation).  A plausible mistake in the
oracle (a dropped term, a swapped channel or index, a wrong sign) fails one of
them.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

W, H = 32, 24
K0 = (30.0, 30.0, 15.5, 11.5)      # fx, fy, cx, cy of the tiny test images


def kin(orc, level=0):
    return orc.level_intr(orc.intr(*K0), level)


# ------------------------------------------------------------------ exp
def test_bexpf_matches_math_exp(orc):
    xs = np.concatenate([np.linspace(-87.0, 0.0, 4001), -np.logspace(-8, 1.9, 500)])
    worst = 0.0
    for x in xs:
        x = float(np.float32(x))
        e = math.exp(x)
        worst = max(worst, abs(orc.bexpf(x) - e) / e)
    assert worst < 4e-7
    assert orc.bexpf(-87.5) == 0.0 and orc.bexpf(-1e4) == 0.0
    assert orc.bexpf(0.0) == 1.0


# ------------------------------------------------------------------ a1 bilateral
def test_bilateral_constant_image_is_identity(orc):
    for v in (300, 1234, 2999):
        img = np.full((H, W), v, np.uint16)
        assert np.array_equal(orc.bilateral(img, 3.0), img)


def test_bilateral_range_gate(orc):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 6000, size=(H, W)).astype(np.uint16)
    out = orc.bilateral(img, 3.0)
    gated = (img > 3000) | (img < 300)
    assert np.all(out[gated] == 0)
    assert np.all(out[~gated] > 0)


def _bilateral_closed_form(img, cx, cy):
    """Output at (cx, cy), restated with math.exp: window weights
    exp(-(d^2 * 0.024691358 + dv^2 * 0.000555556)), dv against the centre
    texel, over the 13x13 window clipped to the image."""
    centre = int(img[cy, cx])
    num = den = 0.0
    for y in range(max(cy - 6, 0), min(cy + 7, H)):
        for x in range(max(cx - 6, 0), min(cx + 7, W)):
            v = int(img[y, x])
            d2 = (x - cx) ** 2 + (y - cy) ** 2
            w = math.exp(-(d2 * 0.024691358 + (v - centre) ** 2 * 0.000555556))
            num += w * v
            den += w
    return num / den


@pytest.mark.parametrize("sx,sy", [(16, 12), (1, 1), (W - 2, H - 1)])
def test_bilateral_spike_closed_form(orc, sx, sy):
    bg, spike = 1500, 1530
    img = np.full((H, W), bg, np.uint16)
    img[sy, sx] = spike
    out = orc.bilateral(img, 3.0)
    for dy in range(-7, 8):
        for dx in range(-7, 8):
            x, y = sx + dx, sy + dy
            if not (0 <= x < W and 0 <= y < H):
                continue
            # the centre texel is gated on its own value (spike is in range)
            ref = _bilateral_closed_form(img, x, y)
            assert abs(int(out[y, x]) - ref) <= 0.5 + 1e-3, (dx, dy, out[y, x], ref)


def test_bilateral_neighbours_not_validity_tested(orc):
    # a neighbour below the 300 mm gate is zeroed as a centre but still enters
    # its neighbours' windows with its range weight
    img = np.full((H, W), 310, np.uint16)
    img[10:15, 19:24] = 270
    out = orc.bilateral(img, 3.0)
    assert np.all(out[10:15, 19:24] == 0)
    ref = _bilateral_closed_form(img, 18, 12)
    assert ref < 310 - 1.5
    assert abs(int(out[12, 18]) - ref) <= 0.5 + 1e-3


# ------------------------------------------------------------------ a2 pyrDown u16
def test_pyrdown_u16_constant_and_ramp(orc):
    img = np.full((H, W), 2345, np.uint16)
    assert np.all(orc.pyrdown_u16(img) == 2345)
    # ramp of 20 mm per column: every window tap is within 90 of the centre and
    # the symmetric weights reproduce the centre value in the interior
    ramp = (1000 + 20 * np.arange(W))[None, :].repeat(H, 0).astype(np.uint16)
    out = orc.pyrdown_u16(ramp)
    for y in range(1, H // 2 - 1):
        for x in range(1, W // 2 - 1):
            assert out[y, x] == ramp[2 * y, 2 * x]


def test_pyrdown_u16_edge_does_not_bleed(orc):
    img = np.full((H, W), 1000, np.uint16)
    img[:, W // 2:] = 1200                      # step of 200 > 90
    out = orc.pyrdown_u16(img)
    assert np.all(out[:, : W // 4] == 1000)
    assert np.all(out[:, W // 4:] == 1200)


def test_pyrdown_u16_brute_force(orc):
    """Clipped 5x5 window, weights w[|dx|] w[|dy|], taps within 90 of the
    centre, truncated quotient -- evaluated independently in float64 (the
    weights are dyadic, so float32 and float64 agree on these sums)."""
    rng = np.random.default_rng(3)
    img = rng.integers(900, 1100, size=(H, W)).astype(np.uint16)
    out = orc.pyrdown_u16(img)
    w = [0.375, 0.25, 0.0625]
    for y in range(H // 2):
        for x in range(W // 2):
            c = int(img[2 * y, 2 * x])
            s = ws = 0.0
            for yy in range(max(2 * y - 2, 0), min(2 * y + 3, H)):
                for xx in range(max(2 * x - 2, 0), min(2 * x + 3, W)):
                    v = int(img[yy, xx])
                    if abs(v - c) < 90:
                        wt = w[abs(xx - 2 * x)] * w[abs(yy - 2 * y)]
                        s += wt * v
                        ws += wt
            assert abs(int(out[y, x]) - int(s / ws)) <= 1, (x, y)


# ------------------------------------------------------------------ a3/a4 maps
def test_vmap_closed_form_and_invalid(orc):
    d = np.full((H, W), 2000, np.uint16)
    d[3, 4] = 0
    d[5, 6] = 21000                                  # beyond the 20 m cutoff
    vm = orc.vmap(d, kin(orc), 20.0)
    fx, fy, cx, cy = K0
    u = np.arange(W)[None, :]
    v = np.arange(H)[:, None]
    ok = d > 0
    ok[5, 6] = False
    assert np.isnan(vm[0][~ok]).all()
    np.testing.assert_allclose(vm[2][ok], 2.0, rtol=0, atol=0)
    np.testing.assert_allclose(vm[0][ok], (2.0 * (u - cx) / fx + 0 * v)[ok], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(vm[1][ok], (2.0 * (v - cy) / fy + 0 * u)[ok], rtol=1e-6, atol=1e-6)


def test_nmap_of_planes(orc):
    # fronto-parallel plane: the normal is the constant viewing axis
    d = np.full((H, W), 1500, np.uint16)
    nm = orc.nmap(orc.vmap(d, kin(orc), 20.0))
    inner = (slice(0, H - 1), slice(0, W - 1))
    np.testing.assert_allclose(np.abs(nm[2][inner]), 1.0, atol=1e-6)
    np.testing.assert_allclose(nm[0][inner], 0.0, atol=1e-6)
    assert np.isnan(nm[0][H - 1, :]).all() and np.isnan(nm[0][:, W - 1]).all()
    # tilted plane z = 1.5 + 0.5 x: normal parallel to (-0.5, 0, 1)
    fx, fy, cx, cy = K0
    vm = np.zeros((3, H, W), np.float32)
    u = np.arange(W)[None, :].repeat(H, 0)
    v = np.arange(H)[:, None].repeat(W, 1)
    a = (u - cx) / fx
    z = 1.5 / (1.0 - 0.5 * a)
    vm[0], vm[1], vm[2] = a * z, (v - cy) / fy * z, z
    nm = orc.nmap(vm)
    ref = np.array([-0.5, 0.0, 1.0]) / math.sqrt(1.25)
    n = nm[:, : H - 1, : W - 1].reshape(3, -1)
    cos = np.abs(ref @ n)
    assert cos.min() > 1 - 1e-5
    # a NaN vertex invalidates itself and the pixels whose stencil reads it
    vm[0, 5, 5] = np.nan
    nm = orc.nmap(vm)
    assert np.isnan(nm[0, 5, 5]) and np.isnan(nm[0, 4, 5]) and np.isnan(nm[0, 5, 4])


def test_copy_maps_and_vertices_to_depth(orc):
    rng = np.random.default_rng(5)
    v4 = rng.uniform(-1, 7, size=(H, W, 4)).astype(np.float32)
    n4 = rng.uniform(-1, 1, size=(H, W, 4)).astype(np.float32)
    v4[2, 3, 2] = 0.0
    vo, no = orc.copy_maps(v4, n4)
    zero = v4[..., 2] == 0
    assert np.isnan(vo[:, zero]).all() and np.isnan(no[:, zero]).all()
    for c in range(3):
        np.testing.assert_array_equal(vo[c][~zero], v4[..., c][~zero])
        np.testing.assert_array_equal(no[c][~zero], n4[..., c][~zero])
    dz = orc.vertices_to_depth(v4, 6.0)
    z = v4[..., 2]
    bad = (z > 6.0) | (z <= 0.0)
    assert np.isnan(dz[bad]).all()
    np.testing.assert_array_equal(dz[~bad], z[~bad])


def test_resize_map_means_and_nan(orc):
    rng = np.random.default_rng(7)
    m = rng.uniform(0.5, 2.0, size=(3, H, W)).astype(np.float32)
    m[0, 4, 6] = np.nan
    r = orc.resize_map(m, False)
    mean = m.reshape(3, H // 2, 2, W // 2, 2).astype(np.float64).mean(axis=(2, 4))
    bad = np.isnan(mean[0])
    assert bad.sum() == 1 and np.isnan(r[:, bad]).all()
    np.testing.assert_allclose(r[:, ~bad], mean[:, ~bad], rtol=1e-6)
    rn = orc.resize_map(m, True)
    nrm = mean / np.linalg.norm(mean, axis=0, keepdims=True)
    np.testing.assert_allclose(rn[:, ~bad], nrm[:, ~bad], rtol=1e-5, atol=1e-6)


def test_transform_maps_rotation_about_z(orc):
    rng = np.random.default_rng(9)
    vm = rng.uniform(-1, 1, size=(3, H, W)).astype(np.float32)
    nm = rng.uniform(-1, 1, size=(3, H, W)).astype(np.float32)
    vm[0, 0, 0] = np.nan
    R = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    t = np.array([0.5, -0.25, 2.0], np.float32)
    v2, n2 = orc.transform_maps(vm, nm, R, t)
    np.testing.assert_allclose(v2[0].ravel()[1:], (-vm[1] + 0.5).ravel()[1:], atol=1e-6)
    np.testing.assert_allclose(v2[1].ravel()[1:], (vm[0] - 0.25).ravel()[1:], atol=1e-6)
    np.testing.assert_allclose(v2[2].ravel()[1:], (vm[2] + 2.0).ravel()[1:], atol=1e-6)
    np.testing.assert_allclose(n2[0], -nm[1], atol=1e-6)
    assert np.isnan(v2[0, 0, 0])


# ------------------------------------------------------------------ a9 / a11 Gaussian pyramids
def _binomial_interior(img):
    k = np.array([1, 4, 6, 4, 1], np.float64)
    K = np.outer(k, k) / 256.0
    out = {}
    for y in range(1, img.shape[0] // 2 - 1):
        for x in range(1, img.shape[1] // 2 - 1):
            out[(y, x)] = float((img[2 * y - 2:2 * y + 3, 2 * x - 2:2 * x + 3] * K).sum())
    return out


def test_pyrdown_f32_interior_is_binomial(orc):
    rng = np.random.default_rng(11)
    img = rng.uniform(0.5, 4.0, size=(H, W)).astype(np.float32)
    out = orc.pyrdown_f32(img)
    for (y, x), ref in _binomial_interior(img.astype(np.float64)).items():
        assert abs(out[y, x] - ref) <= 2e-6 * ref
    c = np.full((H, W), 1.75, np.float32)
    np.testing.assert_allclose(orc.pyrdown_f32(c), 1.75, rtol=1e-7)
    allnan = np.full((H, W), np.nan, np.float32)
    assert np.isnan(orc.pyrdown_f32(allnan)).all()


def test_pyrdown_f32_border_uses_last_column_exclusion(orc):
    # the window's upper bound min(2x+3, cols-1) leaves the last column out:
    # a NaN-free image whose last column is huge must not see it
    img = np.ones((H, W), np.float32)
    img[:, W - 1] = 1e6
    img[H - 1, :] = 1e6
    out = orc.pyrdown_f32(img)
    np.testing.assert_allclose(out, 1.0, rtol=1e-7)


def test_pyrdown_u8_constant_and_zero_exclusion(orc):
    img = np.full((H, W), 77, np.uint8)
    assert np.all(orc.pyrdown_u8(img) == 77)
    chk = np.where((np.add.outer(np.arange(H), np.arange(W)) % 2) == 0, 90, 0).astype(np.uint8)
    out = orc.pyrdown_u8(chk)
    assert np.all(out == 90)          # zeros are not counted
    assert np.all(orc.pyrdown_u8(np.zeros((H, W), np.uint8)) == 0)


def test_pyrdown_u8_interior_is_binomial(orc):
    rng = np.random.default_rng(13)
    img = rng.integers(1, 256, size=(H, W)).astype(np.uint8)
    out = orc.pyrdown_u8(img)
    for (y, x), ref in _binomial_interior(img.astype(np.float64)).items():
        assert int(out[y, x]) == int(ref + 1e-4) or abs(out[y, x] - ref) < 1e-3


# ------------------------------------------------------------------ a10 intensity
def test_intensity_channel_weights(orc):
    px = np.array([[[100, 0, 0], [0, 100, 0], [0, 0, 100], [200, 200, 200], [0, 0, 0]]], np.uint8)
    out = orc.intensity(px)[0]
    assert list(out[:3]) == [11, 29, 58]          # 0.114, 0.299, 0.587 on bytes 0, 1, 2
    assert out[3] in (199, 200) and out[4] == 0
    # four-byte texels use the same three bytes and ignore the fourth
    px4 = np.concatenate([px, np.full(px.shape[:2] + (1,), 255, np.uint8)], axis=2)
    assert np.array_equal(orc.intensity(px4), orc.intensity(px))


def test_intensity_monotone_exhaustive(orc):
    v = np.arange(0, 256, 15, dtype=np.uint8)
    g = np.stack(np.meshgrid(v, v, v, indexing="ij"), -1).reshape(1, -1, 3)
    out = orc.intensity(g).reshape(len(v), len(v), len(v)).astype(int)
    assert (np.diff(out, axis=0) >= 0).all() and (np.diff(out, axis=1) >= 0).all()
    assert (np.diff(out, axis=2) >= 0).all()
    exact = 0.114 * g[0, :, 0] + 0.299 * g[0, :, 1] + 0.587 * g[0, :, 2]
    assert np.all(np.abs(out.ravel() - exact) < 1.0 + 1e-4)


# ------------------------------------------------------------------ a12 derivatives
def test_derivative_of_ramps(orc):
    sx = 2 * 0.52201 + 0.79451
    for b in (1, 3, 7):
        ramp_x = (b * np.arange(W))[None, :].repeat(H, 0).astype(np.uint8)
        dx, dy = orc.derivative(ramp_x)
        inner = (slice(1, H - 1), slice(1, W - 1))
        assert np.all(dx[inner] == int(2 * b * sx))
        assert np.all(dy[inner] == 0)
        ramp_y = (b * np.arange(H))[:, None].repeat(W, 1).astype(np.uint8)
        dx, dy = orc.derivative(ramp_y)
        assert np.all(dy[inner] == int(2 * b * sx)) and np.all(dx[inner] == 0)


def test_derivative_interior_is_convolution(orc):
    rng = np.random.default_rng(17)
    img = rng.integers(0, 256, size=(H, W)).astype(np.uint8)
    dx, dy = orc.derivative(img)
    k = np.array([[0.52201, 0.0, -0.52201], [0.79451, 0.0, -0.79451], [0.52201, 0.0, -0.52201]])
    f = img.astype(np.float64)
    for y in range(1, H - 1):
        for x in range(1, W - 1):
            win = f[y - 1:y + 2, x - 1:x + 2]
            gx = (win[:, ::-1] * k).sum()            # convolution: flipped taps
            gy = (win[::-1, :] * k.T).sum()
            assert abs(dx[y, x] - math.trunc(gx)) <= 1 and abs(dy[y, x] - math.trunc(gy)) <= 1


def test_derivative_border_weight_shift(orc):
    # The tap index counts down from 8 over the in-bounds taps only, so at the
    # border the n surviving taps (row-major) take weights k[8], ..., k[9-n].
    # On a constant image the result is v * (sum of the last n weights).
    kx = [0.52201, 0.0, -0.52201, 0.79451, 0.0, -0.79451, 0.52201, 0.0, -0.52201]
    ky = [0.52201, 0.79451, 0.52201, 0.0, 0.0, 0.0, -0.52201, -0.79451, -0.52201]
    v = 100
    img = np.full((H, W), v, np.uint8)
    dx, dy = orc.derivative(img)
    for y in range(H):
        for x in range(W):
            n = (min(y + 1, H - 1) - max(y - 1, 0) + 1) * (min(x + 1, W - 1) - max(x - 1, 0) + 1)
            ex = math.trunc(v * sum(kx[9 - n:]))
            ey = math.trunc(v * sum(ky[9 - n:]))
            assert abs(dx[y, x] - ex) <= 1 and abs(dy[y, x] - ey) <= 1, (x, y, n, dx[y, x], dy[y, x])
    assert dx[0, 0] == -79 and dy[0, 5] == -183


# ------------------------------------------------------------------ a13 point cloud
def test_point_cloud_closed_form(orc):
    z = np.full((H, W), 2.5, np.float32)
    z[1, 1] = np.nan
    pc = orc.point_cloud(z, kin(orc, 1))
    k1 = kin(orc, 1)
    assert abs(k1.fx - K0[0] / 2) < 1e-7 and abs(k1.cx - K0[2] / 2) < 1e-7
    u = np.arange(W)[None, :]
    v = np.arange(H)[:, None]
    ok = ~np.isnan(z)
    np.testing.assert_allclose(pc[0][ok], ((u - k1.cx) * 2.5 / k1.fx + 0 * v)[ok], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(pc[1][ok], ((v - k1.cy) * 2.5 / k1.fy + 0 * u)[ok], rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(pc[2][ok], 2.5)


def test_gauss_pyramids_border_weight_shift(orc):
    """pyrDownKernelGaussF / IntensityGauss (utils.cu:332-363, 470-500) index
    the weights as (ty - cy - 1) * 5 + (tx - cx - 1) over rows cy < ty =
    min(2y + 3, rows - 1).  On the first output row that gives source rows
    0, 1, 2 the weights 6, 4, 1 (a correctly clipped Gaussian); on the last
    one (2y = rows - 2) source rows rows-4, rows-3, rows-2 get 6, 4, 1 -- the
    weights shift away from the centre.  A ramp in y exposes both."""
    ramp = np.arange(H, dtype=np.float32)[:, None].repeat(W, 1) + 10.0
    out = orc.pyrdown_f32(ramp)
    top = (6 * 10.0 + 4 * 11.0 + 1 * 12.0) / 11.0
    bottom = (6 * (H - 4) + 4 * (H - 3) + 1 * (H - 2)) / 11.0 + 10.0
    np.testing.assert_allclose(out[0, 1:-1], top, rtol=1e-6)
    np.testing.assert_allclose(out[-1, 1:-1], bottom, rtol=1e-6)
    # interior rows stay centred
    np.testing.assert_allclose(out[3, 1:-1], 16.0, rtol=1e-6)
    out8 = orc.pyrdown_u8(ramp.astype(np.uint8))
    assert np.all(out8[0, 1:-1] == int(top)) and np.all(out8[-1, 1:-1] == int(bottom))
