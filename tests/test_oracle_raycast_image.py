"""Pins of the oracle's ray caster and image operations (CPU only): analytic
surfaces, a numpy brute force of the z-range tiles, and closed forms of the
image pyramid operations (OpenCV's pyrDown restated: parity unpinned against
OpenCV itself, which is absent; pinned here by its defining properties)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


def plane_map(D=1.2, W=64, H=48, vs=0.01, mu=0.03, frames=2):
    K = np.array([60.0, 60.0, 31.5, 23.5])
    om = po.OracleMap(vs, mu)
    depth = np.full((H, W), D, np.float32)
    for _ in range(frames):
        om.fuse(depth, K, np.eye(4))
    return om, K, W, H


def test_raycast_plane_hits_the_analytic_surface():
    om, K, W, H = plane_map()
    v, zmin, zmax = om.raycast(W, H, K, np.eye(4))
    hit = v[..., 3] > 0
    assert hit.mean() > 0.9
    assert np.abs(v[..., 2][hit] - 1.2).max() <= 0.5 * 0.01  # within half a voxel of z = 1.2
    # pixels the ray caster reports as misses are written w = -1
    assert set(np.unique(v[..., 3])) <= {-1.0, 1.0}


def test_raycast_sphere_hits_the_analytic_surface():
    W, H = 80, 64
    K = np.array([70.0, 70.0, 39.5, 31.5])
    c, r = np.array([0.05, -0.03, 1.0]), 0.25
    ys, xs = np.mgrid[0:H, 0:W]
    d = np.stack([(xs - K[2]) / K[0], (ys - K[3]) / K[1], np.ones_like(xs, float)], -1)
    # z-depth of the first ray/sphere intersection (ray parameter with d_z = 1)
    A = (d * d).sum(-1)
    B = -2 * (d * c).sum(-1)
    C = (c * c).sum() - r * r
    disc = B * B - 4 * A * C
    z = np.where(disc > 0, (-B - np.sqrt(np.maximum(disc, 0))) / (2 * A), 0.0).astype(np.float32)
    om = po.OracleMap(0.005, 0.02)
    for _ in range(2):
        om.fuse(z, K, np.eye(4))
    v, _, _ = om.raycast(W, H, K, np.eye(4))
    hit = v[..., 3] > 0
    assert hit.sum() > 0.8 * (z > 0).sum()
    dist = np.abs(np.linalg.norm(v[..., :3][hit] - c, axis=1) - r)
    assert np.percentile(dist, 99) <= 0.5 * 0.005


def test_raycast_refuses_sizes_the_zrange_grid_does_not_cover():
    # (W/8)*(H/8) cells: the reference reads cell (x/8, y/8) for every pixel,
    # past the grid for the last H % 8 rows; the oracle refuses such sizes,
    # as the library does, instead of reading there.
    om, K, _, _ = plane_map()
    for W, H in ((64, 60), (60, 48), (4, 8)):
        with pytest.raises(ValueError):
            om.raycast(W, H, K, np.eye(4))
    with pytest.raises(ValueError):
        po.OraclePipeline(80, 60, K, 0.01, 0.03, 1, (1,), po.MODE_DEPTH)


def test_zrange_tiles_equal_numpy_brute_force():
    rng = np.random.default_rng(11)
    W, H = 64, 48
    K = np.array([50.0, 50.0, 31.5, 23.5])
    depth = rng.uniform(0.5, 2.0, (H, W)).astype(np.float32)
    om = po.OracleMap(0.04, 0.12)
    om.fuse(depth, K, np.eye(4))
    T = po.se3_exp(np.array([0.03, -0.02, 0.05, 0.02, -0.03, 0.01]))
    _, zmin, zmax = om.raycast(W, H, K, T)
    vis = om.visible(W, H, K, T)
    Ti = np.linalg.inv(T)
    R, t = Ti[:3, :3].astype(np.float32), Ti[:3, 3].astype(np.float32)
    fx, fy, cx, cy = (f32(v) for v in K)
    cols, rows = W // 8, H // 8
    ezmin = np.full((rows, cols), 100.0, np.float32)
    ezmax = np.zeros((rows, cols), np.float32)
    scale = f32(0.04) * f32(8)
    for b in vis:
        ul, lr = [cols, rows], [-1, -1]
        zr = [f32(3.0), f32(0.1)]
        for c in range(8):
            p = [f32(b[0] + (c & 1)) * scale, f32(b[1] + ((c >> 1) & 1)) * scale, f32(b[2] + ((c >> 2) & 1)) * scale]
            q = [R[i, 0] * p[0] + R[i, 1] * p[1] + R[i, 2] * p[2] + t[i] for i in range(3)]
            pu, pv = (fx * q[0] / q[2] + cx) / f32(8), (fy * q[1] / q[2] + cy) / f32(8)
            ul[0] = min(ul[0], math.floor(pu)); lr[0] = max(lr[0], math.ceil(pu))
            ul[1] = min(ul[1], math.floor(pv)); lr[1] = max(lr[1], math.ceil(pv))
            zr[0] = min(zr[0], q[2]); zr[1] = max(zr[1], q[2])
        ul = [max(ul[0], 0), max(ul[1], 0)]
        lr = [min(lr[0], cols - 1), min(lr[1], rows - 1)]
        if ul[0] > lr[0] or ul[1] > lr[1]:
            continue
        zr[0] = max(zr[0], f32(0.1))
        ezmin[ul[1]:lr[1] + 1, ul[0]:lr[0] + 1] = np.minimum(ezmin[ul[1]:lr[1] + 1, ul[0]:lr[0] + 1], zr[0])
        ezmax[ul[1]:lr[1] + 1, ul[0]:lr[0] + 1] = np.maximum(ezmax[ul[1]:lr[1] + 1, ul[0]:lr[0] + 1], zr[1])
    np.testing.assert_array_equal(zmin.astype(np.float32), ezmin)
    np.testing.assert_array_equal(zmax.astype(np.float32), ezmax)


def np_pyrdown(img):
    """5x5 binomial filter with numpy's 'reflect' padding (= BORDER_REFLECT_101),
    then every second pixel, in float64."""
    k = np.array([1, 4, 6, 4, 1], float) / 16.0
    p = np.pad(img.astype(float), 2, mode="reflect")
    v = sum(k[i] * p[i:i + img.shape[0], :] for i in range(5))
    h = sum(k[j] * v[:, j:j + img.shape[1]] for j in range(5))
    return h[::2, ::2]


def test_pyrdown_matches_binomial_filter_with_reflect101():
    rng = np.random.default_rng(12)
    img = rng.uniform(0, 255, (30, 44)).astype(np.float32)
    np.testing.assert_allclose(po.pyrdown_gauss(img, 64), np_pyrdown(img), rtol=0, atol=1e-9)
    np.testing.assert_allclose(po.pyrdown_gauss(img, 32), np_pyrdown(img), rtol=0, atol=1e-3)


def test_pyrdown_preserves_constants_and_ramps():
    c = np.full((24, 32), 37.25, np.float32)
    assert np.all(po.pyrdown_gauss(c) == 37.25)
    ys, xs = np.mgrid[0:24, 0:32]
    ramp = (0.5 * xs + 0.25 * ys + 3.0).astype(np.float32)
    out = po.pyrdown_gauss(ramp, 64)
    inner = out[1:-1, 1:-1]
    yi, xi = np.mgrid[1:11, 1:15]
    np.testing.assert_allclose(inner, 0.5 * 2 * xi + 0.25 * 2 * yi + 3.0, atol=1e-12)


def test_gradient_closed_form():
    ys, xs = np.mgrid[0:20, 0:30]
    img = (2.0 * xs - 3.0 * ys + 7.0).astype(np.float32)
    gx, gy = po.gradient(img)
    assert np.all(gx[1:-1, 1:-1] == 2.0) and np.all(gy[1:-1, 1:-1] == -3.0)
    assert np.all(gx[0] == 0) and np.all(gx[-1] == 0) and np.all(gx[:, 0] == 0) and np.all(gy[:, -1] == 0)


def test_inverse_depth_vertex_and_normal_maps_of_a_plane():
    H, W = 24, 32
    z = np.full((H, W), 1.5, np.float32)
    z[0, 0], z[1, 1], z[2, 2] = 0.1, 9.0, 0.0
    inv = po.depth_to_invdepth(z)
    assert inv[0, 0] == 0 and inv[1, 1] == 0 and inv[2, 2] == 0
    assert inv[5, 5] == f32(1.0) / f32(1.5)
    fx, fy, cx, cy = 30.0, 31.0, 15.5, 11.5
    invfx, invfy = float(f32(1 / fx)), float(f32(1 / fy))
    v = po.vertex_map(inv, invfx, invfy, cx, cy, 64)
    ys, xs = np.mgrid[0:H, 0:W]
    ok = inv > 0
    with np.errstate(divide="ignore"):
        zz = 1.0 / inv.astype(np.float64)  # z = 1/invz as the vertex kernel forms it
    np.testing.assert_allclose(v[..., 2][ok], zz[ok], rtol=1e-15)
    np.testing.assert_allclose(v[..., 0][ok], (zz * (xs - cx) * invfx)[ok], rtol=1e-12)
    np.testing.assert_allclose(v[..., 1][ok], (zz * (ys - cy) * invfy)[ok], rtol=1e-12)
    np.testing.assert_allclose(v[..., 2][ok], 1.5, rtol=1e-7)
    assert np.all(v[..., 3][~ok] == -1)
    n = po.normal_map(v.astype(np.float32), 64)
    good = n[..., 3] > 0
    assert good.sum() > 0.8 * H * W
    np.testing.assert_allclose(n[..., :3][good], np.tile([0.0, 0.0, -1.0], (good.sum(), 1)), atol=1e-6)
    assert np.all(n[-1, :, 3] == -1) and np.all(n[:, -1, 3] == -1)
