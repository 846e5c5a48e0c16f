"""Pins of the oracle's allocation, visibility and integration (CPU only).

The checks restate the reference's arithmetic independently in numpy
float32 (one IEEE operation per step, same order as VoxelStructUtils.h and
VoxelMap.cu), test a closed form (a fronto-parallel plane), and measure the
storage variants against the FP64 oracle.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from yetanotherslam_amd import synth

f32 = np.float32


def walk(b0, b1):
    """Integer 3D Bresenham of VoxelMap.cu:568-650 (dominant axis first, the
    z branch advancing y before x)."""
    d = [b1[i] - b0[i] for i in range(3)]
    inc = [-1 if v < 0 else 1 for v in d]
    ad = [abs(v) for v in d]
    e2 = [2 * v for v in ad]
    if ad[0] >= ad[1] and ad[0] >= ad[2]:
        a, o1, o2 = 0, 1, 2
    elif ad[1] >= ad[0] and ad[1] >= ad[2]:
        a, o1, o2 = 1, 0, 2
    else:
        a, o1, o2 = 2, 1, 0
    p = list(b0)
    err1, err2 = e2[o1] - 1, e2[o2] - 1
    out = [tuple(p)]
    for _ in range(ad[a]):
        if err1 > 0:
            p[o1] += inc[o1]
            err1 -= e2[a]
        if err2 > 0:
            p[o2] += inc[o2]
            err2 -= e2[a]
        err1 += e2[o1]
        err2 += e2[o2]
        p[a] += inc[a]
        out.append(tuple(p))
    return out


def test_bresenham_walk_properties():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        b0 = tuple(int(v) for v in rng.integers(-20, 20, 3))
        b1 = tuple(int(v) for v in rng.integers(-20, 20, 3))
        w = walk(b0, b1)
        assert w[0] == b0 and w[-1] == b1
        assert len(w) == max(abs(b1[i] - b0[i]) for i in range(3)) + 1
        for a, b in zip(w, w[1:]):
            assert max(abs(a[i] - b[i]) for i in range(3)) == 1  # 26-adjacent steps
        assert len(set(w)) == len(w)


def np_block(x, y, z, K, R, t, vs):
    """VoxelStructUtils.h:97-110 then :28-49 in float32."""
    fx, fy, cx, cy = (f32(v) for v in K)
    invfx, invfy = f32(1.0 / np.float64(fx)), f32(1.0 / np.float64(fy))
    pc = [invfx * (f32(x) - cx) * z, invfy * (f32(y) - cy) * z, z]
    out = []
    for i in range(3):
        w = R[i, 0] * pc[0] + R[i, 1] * pc[1] + R[i, 2] * pc[2] + t[i]
        v = int(math.floor(w / vs))
        out.append(v // 8)
    return tuple(out)


def np_allocate(depth, K, T, vs, mu, dmin=f32(0.1), dmax=f32(3.0)):
    R, t = T[:3, :3].astype(np.float32), T[:3, 3].astype(np.float32)
    vs, mu = f32(vs), f32(mu)
    half = mu * f32(0.5)
    blocks = set()
    H, W = depth.shape
    for y in range(H):
        for x in range(W):
            d = depth[y, x]
            if np.isnan(d) or d < dmin or d > dmax:
                continue
            dn, df = max(dmin, d - half), min(dmax, d + half)
            if dn >= df:
                continue
            blocks.update(walk(np_block(x, y, dn, K, R, t, vs), np_block(x, y, df, K, R, t, vs)))
    return blocks


def random_pose(rng, rot=0.3, trans=0.5):
    w = rng.normal(size=3) * rot
    T = po.se3_exp(np.concatenate([rng.normal(size=3) * trans, w]))
    return T


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_allocation_block_set_equals_numpy_brute_force(seed):
    rng = np.random.default_rng(seed)
    W, H = 24, 18
    K = np.array([30.0, 31.0, 11.5, 8.5])
    depth = rng.uniform(0.05, 3.2, (H, W)).astype(np.float32)
    depth[rng.random((H, W)) < 0.1] = np.nan
    T = random_pose(rng)
    vs, mu = 0.02, 0.07
    om = po.OracleMap(vs, mu)
    om.allocate(depth, K, T)
    got = {tuple(k) for k in om.export(with_voxels=False)}
    ref = np_allocate(depth, K, T, vs, mu)
    assert got == ref and len(ref) > 50


def test_visibility_equals_numpy_brute_force():
    rng = np.random.default_rng(7)
    W, H = 32, 24
    K = np.array([40.0, 40.0, 15.5, 11.5])
    depth = rng.uniform(0.3, 2.5, (H, W)).astype(np.float32)
    om = po.OracleMap(0.03, 0.09)
    om.allocate(depth, K, np.eye(4))
    keys = om.export(with_voxels=False)
    T = random_pose(rng, 0.2, 0.2)
    Ti = np.linalg.inv(T)
    R, t = Ti[:3, :3].astype(np.float32), Ti[:3, 3].astype(np.float32)
    fx, fy, cx, cy = (f32(v) for v in K)
    scale = f32(0.03) * f32(8)
    ref = set()
    for b in keys:
        for c in range(8):
            p = [f32(b[0] + (c & 1)) * scale, f32(b[1] + ((c >> 1) & 1)) * scale, f32(b[2] + ((c >> 2) & 1)) * scale]
            q = [R[i, 0] * p[0] + R[i, 1] * p[1] + R[i, 2] * p[2] + t[i] for i in range(3)]
            with np.errstate(all="ignore"):
                u, v = fx * q[0] / q[2] + cx, fy * q[1] / q[2] + cy
            if not (u < 0 or v < 0 or u > W - 1 or v > H - 1 or q[2] < f32(0.1) or q[2] > f32(3.0)):
                ref.add(tuple(int(x) for x in b))
                break
    got = {tuple(k) for k in om.visible(W, H, K, T)}
    assert got == ref and 0 < len(ref) < len(keys)


@pytest.mark.parametrize("voxfmt", [po.VOX_F32, po.VOX_REF, po.VOX_F16])
def test_integration_closed_form_plane(voxfmt):
    """A fronto-parallel plane at depth D seen from the identity pose: every
    voxel at depth z that projects into the image holds
    min(1, (D - z) / mu) (quantised by the storage), weight = frames."""
    W, H = 32, 24
    K = np.array([30.0, 30.0, 15.5, 11.5])
    D, vs, mu = f32(1.0), 0.01, 0.03
    depth = np.full((H, W), D, np.float32)
    om = po.OracleMap(vs, mu, voxfmt=voxfmt)
    for _ in range(3):
        om.fuse(depth, K, np.eye(4))
    keys, sdf, wt = om.export()
    fx, fy, cx, cy = (f32(v) for v in K)
    n_checked = 0
    for b, s_blk, w_blk in zip(keys, sdf, wt):
        for l in range(512):
            lx, ly, lz = l & 7, (l >> 3) & 7, l >> 6
            x, y, z = (f32(b[0] * 8 + lx) * f32(vs), f32(b[1] * 8 + ly) * f32(vs), f32(b[2] * 8 + lz) * f32(vs))
            u = math.floor(fx * x / z + cx + f32(0.5))
            v = math.floor(fy * y / z + cy + f32(0.5))
            inside = 0 <= u <= W - 1 and 0 <= v <= H - 1 and D - z >= -f32(mu)
            if not inside:
                assert w_blk[l] == 0
                continue
            e = min(f32(1.0), (D - z) / f32(mu))
            if voxfmt == po.VOX_REF:
                e = f32(np.trunc(e * f32(32767.0)) / f32(32767.0))
            elif voxfmt == po.VOX_F16:
                e = f32(np.float16(e))
            assert w_blk[l] == 3
            assert abs(s_blk[l] - e) <= 2e-6, (b, l, s_blk[l], e)
            n_checked += 1
    assert n_checked > 1000


def test_storage_variants_against_fp64_tsdf():
    """TSDF RMS error (in voxels) of the three storage variants against the
    FP64 oracle on four frames of the synthetic sequence at 8 mm."""
    p = synth.params(width=320, height=240, theta0=math.pi / 4)
    K = synth.intrinsics(p)
    gt = synth.relative_poses(p, 4)
    vs, mu = 0.008, 0.024
    maps = {k: po.OracleMap(vs, mu, precision=32, voxfmt=k) for k in (po.VOX_F32, po.VOX_F16, po.VOX_REF)}
    m64 = po.OracleMap(vs, mu, precision=64)
    for t in range(4):
        z = po.depth_u16_to_f32(synth.frame(p, t)[0])
        for m in list(maps.values()) + [m64]:
            m.fuse(z, K, gt[t])
    k64, s64, w64 = m64.export()
    idx64 = {tuple(k): i for i, k in enumerate(k64)}
    for fmt, m in maps.items():
        k, s, w = m.export()
        # block sets may differ at block boundaries between precisions
        common = [(i, idx64[tuple(kk)]) for i, kk in enumerate(k) if tuple(kk) in idx64]
        assert len(common) > 0.99 * max(len(k), len(k64))
        a = np.array([c[0] for c in common])
        b = np.array([c[1] for c in common])
        s, w, s2, w2 = s[a], w[a], s64[b], w64[b]
        both = (w > 0) & (w2 > 0)
        assert ((w > 0) == (w2 > 0)).mean() > 0.999
        rms_vox = np.sqrt(np.mean((s[both] - s2[both]) ** 2)) * mu / vs
        assert rms_vox < 0.1, (fmt, rms_vox)


def voxel_dict(keys, sdf, wt):
    out = {}
    for b, s_blk, w_blk in zip(keys, sdf, wt):
        for l in range(512):
            if w_blk[l] > 0:
                out[(b[0] * 8 + (l & 7), b[1] * 8 + ((l >> 3) & 7), b[2] * 8 + (l >> 6))] = (s_blk[l], w_blk[l])
    return out


def test_map_merge_by_a_whole_voxel_shift_is_a_copy():
    """Map-to-map fusion (VoxelMap.cu:400-449) into an empty map whose pose is
    the source's moved by 3 voxels along x.  With a power-of-two voxel size
    every coordinate is exact, so each destination voxel u samples the source
    exactly at u + 3 (trilinear weights 1, 0): it holds that source voxel when
    all 8 trilinear corners are observed and is unobserved otherwise, and the
    destination blocks are those of the shifted observed source voxels."""
    vs = 2.0 ** -7
    W, H = 32, 24
    K = np.array([30.0, 30.0, 15.5, 11.5])
    rng = np.random.default_rng(3)
    depth = (0.5 + 0.05 * rng.random((H, W))).astype(np.float32)
    src = po.OracleMap(vs, 3 * vs, voxfmt=po.VOX_F32)
    src.fuse(depth, K, np.eye(4))
    dst = po.OracleMap(vs, 3 * vs, voxfmt=po.VOX_F32)
    Tdst = np.eye(4)
    Tdst[0, 3] = 3 * vs
    dst.fuse_map(src, Tdst, np.eye(4))
    S = voxel_dict(*src.export())
    # blocks of the observed source voxels moved by -3 voxels
    want_blocks = {(po.lib().orc_vox_to_block(v[0] - 3), po.lib().orc_vox_to_block(v[1]),
                    po.lib().orc_vox_to_block(v[2])) for v in S}
    keys, sdf, wt = dst.export()
    assert {tuple(k) for k in keys} == want_blocks and len(want_blocks) > 20
    D = voxel_dict(keys, sdf, wt)
    corners = [(dx, dy, dz) for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)]
    n_full = 0
    for b in keys:
        for l in range(512):
            u = (b[0] * 8 + (l & 7), b[1] * 8 + ((l >> 3) & 7), b[2] * 8 + (l >> 6))
            v = (u[0] + 3, u[1], u[2])
            full = all((v[0] + c[0], v[1] + c[1], v[2] + c[2]) in S for c in corners)
            if full:
                assert D[u] == S[v], (u, D.get(u), S[v])
                n_full += 1
            else:
                assert u not in D
    assert n_full > 1000


def test_map_merge_weights_average_two_maps():
    """Merging a map into a copy of itself at the same pose averages equal
    values (the sdf stays, the weights add); a second source seen from a
    moved camera changes sdf only where it overlaps."""
    W, H = 32, 24
    K = np.array([30.0, 30.0, 15.5, 11.5])
    depth = np.full((H, W), 0.8, np.float32)
    a = po.OracleMap(0.01, 0.03, voxfmt=po.VOX_F32)
    b = po.OracleMap(0.01, 0.03, voxfmt=po.VOX_F32)
    for m in (a, b):
        m.fuse(depth, K, np.eye(4))
        m.fuse(depth, K, np.eye(4))
    ka, sa, wa = a.export()
    a.fuse_map(b, np.eye(4), np.eye(4))
    ka2, sa2, wa2 = a.export()
    np.testing.assert_array_equal(ka, ka2)
    # interior voxels (8 corners observed in b) get weight 2 + 2 and keep their value
    merged = wa2 > wa
    assert merged.sum() > 1000
    np.testing.assert_array_equal(wa2[merged], 4.0)
    np.testing.assert_allclose(sa2[merged], sa[merged], rtol=0, atol=1e-6)
    assert np.all(wa2[~merged] == wa[~merged]) and np.all(sa2[~merged] == sa[~merged])


@pytest.mark.parametrize("voxfmt", [po.VOX_F32, po.VOX_F16, po.VOX_REF], ids=["f32", "f16u16", "i16u8"])
def test_import_rebuilds_an_exported_volume(voxfmt):
    """import_blocks (test set-up for meshing a volume built on the device)
    restores exactly what export gives, in every storage: the same blocks and
    voxels, and so the same mesh."""
    p = synth.params(width=160, height=120)
    K = synth.intrinsics(p)
    gt = synth.relative_poses(p, 2)
    src = po.OracleMap(0.01, 0.03, precision=32, voxfmt=voxfmt)
    for t in range(2):
        src.fuse(po.depth_u16_to_f32(synth.frame(p, t)[0]), K, gt[t])
    k, s, w = src.export()
    dst = po.OracleMap(0.01, 0.03, precision=32, voxfmt=voxfmt)
    dst.import_blocks(k, s, w)
    k2, s2, w2 = dst.export()
    np.testing.assert_array_equal(k2, k)
    np.testing.assert_array_equal(s2, s)
    np.testing.assert_array_equal(w2, w)
    m1, m2 = src.mesh(), dst.mesh()
    assert len(k) > 500 and len(m1["tris"]) > 1000
    for name in ("keys", "cases", "counts", "tris", "normals"):
        np.testing.assert_array_equal(m2[name], m1[name])
