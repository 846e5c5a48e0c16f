"""GPU parity: every HIP kernel through the C ABI against the scalar oracle.

Bar (see each test): block sets, TSDF storage, z-range maps, ray-cast vertex
maps, image pyramids and marching-cubes cases/triangles are bit-exact; the
tracker's reduced sums are within FP32 reduction-order tolerance of the
oracle's exact sums, and poses within 1e-4 m / 1e-4 rad.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from yetanotherslam_amd import native
from yetanotherslam_amd import rgbdfusion as rf
from yetanotherslam_amd import synth

from helpers import assert_bits_equal, frames, pose_err, sums_tolerance, to_dev

pytestmark = pytest.mark.gpu

FMTS = [(rf.VOXEL_I16_U8, po.VOX_REF), (rf.VOXEL_F32_F32, po.VOX_F32), (rf.VOXEL_F16_U16, po.VOX_F16)]


def fuse_both(fmt_pair, n=3, width=320, height=240, voxel=0.008, trunc=0.024, buckets=1 << 16, blocks=1 << 15):
    gfmt, ofmt = fmt_pair
    p, K, fr = frames(n, width, height)
    gt = synth.relative_poses(p, n)
    om = po.OracleMap(voxel, trunc, precision=32, voxfmt=ofmt)
    gm = rf.MapStruct(K.astype(np.float32), buckets, blocks, voxel, trunc, voxel_format=gfmt)
    for t in range(n):
        z = po.depth_u16_to_f32(fr[t][0])
        om.fuse(z, K, gt[t])
        gm.Fuse(to_dev(z), gt[t])
    torch.cuda.synchronize()
    return p, K, fr, gt, om, gm


@pytest.mark.parametrize("fmt", FMTS, ids=["i16u8", "f32", "f16u16"])
def test_fuse_block_set_and_tsdf_bit_exact(fmt):
    _, _, _, _, om, gm = fuse_both(fmt)
    ko, so, wo = om.export()
    kg, sg, wg = gm.export_sorted()
    assert gm.stats().flags == 0
    assert len(ko) > 100
    np.testing.assert_array_equal(ko, kg)
    assert_bits_equal(so, sg, "sdf")
    assert_bits_equal(wo, wg, "weight")


def test_fuse_c1_full_resolution_bit_exact():
    """640x480, 8 mm voxels, truncation 3 voxels (the plain-method configuration)."""
    _, _, _, _, om, gm = fuse_both(FMTS[1], n=2, width=640, height=480, buckets=1 << 20, blocks=1 << 17)
    ko, so, wo = om.export()
    kg, sg, wg = gm.export_sorted()
    np.testing.assert_array_equal(ko, kg)
    assert_bits_equal(so, sg, "sdf")
    assert_bits_equal(wo, wg, "weight")


def test_visible_block_set_matches():
    """CheckNumVisibleBlocks (VoxelMap.cu:788-823): the count and the set of
    visible blocks equal the oracle's (the reference's list order is
    nondeterministic, so sets are compared)."""
    p, K, fr, gt, om, gm = fuse_both(FMTS[1])
    for T in (gt[0], gt[2], np.linalg.inv(gt[2]) @ gt[0]):
        ko = om.visible(p.width, p.height, K, T)
        n_gpu = gm.CheckNumVisibleBlocks(p.width, p.height, T)
        kg = gm.visible_blocks()
        assert len(ko) == n_gpu == len(kg) > 100
        assert {tuple(k) for k in kg} == {tuple(k) for k in ko}


def test_raycast_bit_exact():
    p, K, fr, gt, om, gm = fuse_both(FMTS[1])
    rc = rf.RayTraceEngine(p.width, p.height, K.astype(np.float32))
    T = gt[2]
    rc.RayTrace(gm, T)
    vg = rc.GetVMap()
    zmin_g, zmax_g = rc.GetZRange()
    vo, zmin_o, zmax_o = om.raycast(p.width, p.height, K, T)
    assert_bits_equal(zmin_o, zmin_g, "zmin")
    assert_bits_equal(zmax_o, zmax_g, "zmax")
    np.testing.assert_array_equal(vo[..., 3], vg[..., 3])
    valid = vo[..., 3] > 0
    assert valid.mean() > 0.3
    assert_bits_equal(vo[valid][:, :3], vg[valid][:, :3], "vertex")


def test_raycast_bit_exact_with_partial_tiles():
    """200x136 is a multiple of 8 but not of 16, so the march's last column
    and row of 16x16 tiles lie half outside the image (their outside 8x8
    cells do nothing): z-ranges and vertex maps still equal the oracle's."""
    p, K, fr, gt, om, gm = fuse_both(FMTS[1], width=200, height=136)
    rc = rf.RayTraceEngine(p.width, p.height, K.astype(np.float32))
    T = gt[2]
    rc.RayTrace(gm, T)
    vg = rc.GetVMap()
    zmin_g, zmax_g = rc.GetZRange()
    vo, zmin_o, zmax_o = om.raycast(p.width, p.height, K, T)
    assert vg.shape == (136, 200, 4) and zmin_g.shape == (17, 25)
    assert_bits_equal(zmin_o, zmin_g, "zmin")
    assert_bits_equal(zmax_o, zmax_g, "zmax")
    np.testing.assert_array_equal(vo[..., 3], vg[..., 3])
    valid = vo[..., 3] > 0
    assert valid[:, 192:].any() and valid[128:, :].any()
    assert_bits_equal(vo[valid][:, :3], vg[valid][:, :3], "vertex")


def test_nothing_to_see_gives_misses_and_no_triangles():
    """A map with no blocks, and a populated map seen from where none of its
    blocks is in view: every pixel a miss (w = -1 and x = NaN, as the
    reference writes first), every z-range cell at its initial 100 / 0, as
    the oracle's; marching cubes of the empty map gives no triangles."""
    p, K, fr = frames(1)
    W, H, K32 = p.width, p.height, K.astype(np.float32)
    empty = rf.MapStruct(K32, 1 << 14, 1 << 12, 0.008, 0.024)
    _, _, _, gt, om, full = fuse_both(FMTS[1], n=1)
    away = np.eye(4)
    away[:3, :3] = np.diag([-1.0, 1.0, -1.0])  # turned round about the y axis
    rc = rf.RayTraceEngine(W, H, K32)
    for m, T in ((empty, np.eye(4)), (full, gt[0] @ away)):
        rc.RayTrace(m, T)
        v = rc.GetVMap()
        zmin, zmax = rc.GetZRange()
        assert (v[..., 3] == -1.0).all() and np.isnan(v[..., 0]).all()
        assert (zmin == 100.0).all() and (zmax == 0.0).all()
    vo, zo_min, zo_max = om.raycast(W, H, K, gt[0] @ away)
    assert (vo[..., 3] == -1.0).all() and (zo_min == 100.0).all() and (zo_max == 0.0).all()
    assert rf.MeshEngine(1000).Meshify(empty)["num_triangles"] == 0


def test_image_ops_bit_exact():
    p, K, fr = frames(1)
    d, rgb, _ = fr[0]
    z = po.depth_u16_to_f32(d)
    g = po.rgb_to_intensity(rgb)
    ctx = rf.default_context()
    lib = rf.native
    W, H = p.width, p.height
    # device conversions from raw frames
    dz = torch.empty((H, W), dtype=torch.float32, device="cuda")
    dg = torch.empty((H, W), dtype=torch.float32, device="cuda")
    du = to_dev(d)
    dc = to_dev(rgb)
    lib.call("yas_depth_u16_to_f32", ctx.h, rf._dev(du), W * 2, rf._dev(dz), W * 4, W, H, 1000.0)
    lib.call("yas_rgb_to_intensity", ctx.h, rf._dev(dc), W * 3, rf._dev(dg), W * 4, W, H, 0)
    # OpenCV's 8-bit fixed-point gray (System.cpp:47-49)
    dg8 = torch.empty((H, W), dtype=torch.float32, device="cuda")
    lib.call("yas_rgb_to_intensity", ctx.h, rf._dev(dc), W * 3, rf._dev(dg8), W * 4, W, H, 1)
    torch.cuda.synchronize()
    assert_bits_equal(z, dz.cpu().numpy(), "depth u16->f32")
    assert_bits_equal(g, dg.cpu().numpy(), "intensity")
    assert_bits_equal(po.rgb_to_intensity(rgb, mode=1), dg8.cpu().numpy(), "8-bit gray")
    # Gaussian pyramid level
    out = torch.empty((H // 2, W // 2), dtype=torch.float32, device="cuda")
    lib.call("yas_pyrdown_gauss", ctx.h, rf._dev(dg), W * 4, W, H, rf._dev(out), (W // 2) * 4)
    torch.cuda.synchronize()
    assert_bits_equal(po.pyrdown_gauss(g), out.cpu().numpy(), "pyrdown")
    # gradients
    gx = torch.empty((H, W), dtype=torch.float32, device="cuda")
    gy = torch.empty_like(gx)
    lib.call("yas_image_gradient", ctx.h, rf._dev(dg), W * 4, rf._dev(gx), rf._dev(gy), W * 4, W, H)
    torch.cuda.synchronize()
    ogx, ogy = po.gradient(g)
    assert_bits_equal(ogx, gx.cpu().numpy(), "gx")
    assert_bits_equal(ogy, gy.cpu().numpy(), "gy")
    # inverse depth, vertex map, normal map
    inv = torch.empty((H, W), dtype=torch.float32, device="cuda")
    lib.call("yas_depth_to_invdepth", ctx.h, rf._dev(dz), W * 4, rf._dev(inv), W * 4, W, H)
    vm = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    nm = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    Kf = K.astype(np.float32)
    lib.call("yas_compute_vertex_map", ctx.h, rf._dev(inv), W * 4, rf._dev(vm), W * 16, W, H, rf._ptr(Kf), 3.0)
    lib.call("yas_compute_normal_map", ctx.h, rf._dev(vm), W * 16, rf._dev(nm), W * 16, W, H)
    torch.cuda.synchronize()
    oinv = po.depth_to_invdepth(z)
    assert_bits_equal(oinv, inv.cpu().numpy(), "invdepth")
    invfx = float(np.float32(1.0 / np.float64(Kf[0])))
    invfy = float(np.float32(1.0 / np.float64(Kf[1])))
    ovm = po.vertex_map(oinv, invfx, invfy, Kf[2], Kf[3])
    assert_bits_equal(ovm, vm.cpu().numpy(), "vertex map")
    onm = po.normal_map(ovm.astype(np.float32))
    assert_bits_equal(onm, nm.cpu().numpy(), "normal map")


def tracker_pair(mode, levels=3, iters=(10, 5, 3), weights=(1.0, 1.0, 100.0, 10.0), width=320, height=240):
    """Oracle and GPU trackers loaded with frames 0 (reference) and 1 (current);
    the reference model is frame 0's own vertex map."""
    p, K, fr = frames(2, width, height)
    z0, z1 = (po.depth_u16_to_f32(fr[i][0]) for i in range(2))
    g0, g1 = (po.rgb_to_intensity(fr[i][1]) for i in range(2))
    Kf = K.astype(np.float32)
    invfx = float(np.float32(1.0 / np.float64(Kf[0])))
    invfy = float(np.float32(1.0 / np.float64(Kf[1])))
    vm = po.vertex_map(po.depth_to_invdepth(z0), invfx, invfy, Kf[2], Kf[3]).astype(np.float32)
    ot = po.OracleTracker(p.width, p.height, K, levels, iters, mode, weights)
    gt_ = rf.RGBDTracking(p.width, p.height, K, levels, iters, mode, weights)
    ot.set_reference_image(g0); ot.set_reference_depth(z0); ot.set_reference_model(vm)
    ot.set_tracking_image(g1); ot.set_tracking_depth(z1)
    gt_.SetReferenceImage(to_dev(g0)); gt_.SetReferenceDepth(to_dev(z0)); gt_.SetReferenceModel(to_dev(vm))
    gt_.SetTrackingImage(to_dev(g1)); gt_.SetTrackingDepth(to_dev(z1))
    torch.cuda.synchronize()
    DT = np.linalg.inv(synth.pose(p, 1)) @ synth.pose(p, 0)
    return p, K, ot, gt_, DT


def test_tracker_pyramids_bit_exact():
    p, K, ot, gt_, _ = tracker_pair(po.MODE_RGBD)
    for lvl in range(3):
        for which in range(10):
            assert_bits_equal(ot.level(which, lvl), gt_.level(which, lvl), "buffer %d level %d" % (which, lvl))


def test_reference_default_five_level_tracker_matches_oracle():
    """The reference's default tracker (NUM_PYR 5, RGBDTracking.h:7;
    iterations {10,5,3,3,3}, RGBDTracking.cu:159; weights 1/1/100/10): all
    five levels of every buffer bit-exact, and the pose within 1e-4 m /
    1e-4 rad of the oracle's, in each mode."""
    for mode in (po.MODE_RGBD, po.MODE_DEPTH, po.MODE_RGB):
        p, K, ot, gt_, DT = tracker_pair(mode, levels=5, iters=(10, 5, 3, 3, 3))
        if mode == po.MODE_RGBD:
            for lvl in range(5):
                for which in range(10):
                    assert_bits_equal(ot.level(which, lvl), gt_.level(which, lvl), "buffer %d level %d" % (which, lvl))
        To, good_o = ot.get_transform(swap=False)
        Tg, st = gt_.GetTransform(swap=False)
        assert good_o and st.good
        et, er = pose_err(Tg, To)
        assert et < 1e-4 and er < 1e-4, (mode, et, er)
        gt_.close()


@pytest.mark.parametrize("kind", [0, 1], ids=["icp", "photometric"])
def test_step_sums_within_reduction_tolerance(kind):
    p, K, ot, gt_, DT = tracker_pair(po.MODE_RGBD)
    for lvl in range(3):
        for T in (np.eye(4), DT):
            so, st_o = ot.step_sums(lvl, T, kind)
            sg, st_g = gt_.step_sums(lvl, T, kind)
            assert so[28] == sg[28], (lvl, so[28], sg[28])  # counts are exact
            ok = sums_tolerance(sg, so)
            assert ok.all(), (kind, lvl, np.argwhere(~ok).ravel(), sg[~ok], so[~ok])


@pytest.mark.parametrize("mode", [po.MODE_DEPTH, po.MODE_RGB, po.MODE_RGBD], ids=["icp", "rgb", "rgbd"])
def test_get_transform_matches_oracle(mode):
    p, K, ot, gt_, DT = tracker_pair(mode)
    To, good_o = ot.get_transform(swap=False)
    Tg, st = gt_.GetTransform(swap=False)
    assert good_o and st.good
    et, er = pose_err(Tg, To)
    assert et < 1e-4 and er < 1e-4, (et, er)


@pytest.mark.parametrize("fmt", FMTS, ids=["i16u8", "f32", "f16u16"])
def test_mesh_cases_and_triangles_bit_exact(fmt):
    _, _, _, _, om, gm = fuse_both(fmt)
    mo = om.mesh()
    me = rf.MeshEngine(2_000_000)
    mg = me.Meshify(gm, with_cases=True)
    keys, counts, cases, tris, normals = rf.canonical_mesh(mg)
    np.testing.assert_array_equal(mo["keys"], keys)
    np.testing.assert_array_equal(mo["cases"], cases)
    np.testing.assert_array_equal(mo["counts"], counts)
    assert mg["num_triangles"] == len(mo["tris"]) > 1000
    assert_bits_equal(mo["tris"], tris, "vertices")
    assert_bits_equal(mo["normals"], normals, "normals")


@pytest.mark.parametrize("fmt", FMTS[1:], ids=["f32", "f16u16"])
def test_mesh_of_a_large_5mm_volume_bit_exact(fmt):
    """Marching cubes over a 5 mm volume fused from 60 frames at 640x480
    (about 30 k blocks, 2.6 M triangles), float32 and float16 sdf: block
    keys, cube cases and per-block triangle counts are bit-exact, and so are
    the vertices and normals."""
    _, _, _, _, om, gm = fuse_both(fmt, n=60, width=640, height=480, voxel=0.005, trunc=0.02, buckets=1 << 20,
                                   blocks=1 << 17)
    mo = om.mesh()
    me = rf.MeshEngine(4_000_000)
    mg = me.Meshify(gm, with_cases=True)
    keys, counts, cases, tris, normals = rf.canonical_mesh(mg)
    print("large mesh:", len(keys), "blocks,", mg["num_triangles"], "triangles")
    assert len(keys) > 20000 and mg["num_triangles"] > 2_000_000
    np.testing.assert_array_equal(mo["keys"], keys)
    np.testing.assert_array_equal(mo["cases"], cases)
    np.testing.assert_array_equal(mo["counts"], counts)
    assert mg["num_triangles"] == len(mo["tris"])
    assert_bits_equal(mo["tris"], tris, "vertices")
    assert_bits_equal(mo["normals"], normals, "normals")


@pytest.mark.parametrize("fmt", FMTS, ids=["i16u8", "f32", "f16u16"])
def test_indexed_mesh_matches_the_soup_and_shares_vertices(fmt):
    """The indexed mesh expands to the reference-form soup (same triangle
    order, corners within 1e-5 m); every triangle edge is used by at most two
    triangles, and most by exactly two (shared vertices, no duplicates)."""
    _, _, _, _, _, gm = fuse_both(fmt)
    me = rf.MeshEngine(2_000_000)
    soup = me.Meshify(gm)
    ix = me.MeshifyIndexed(gm)
    V, T = ix["vertices"], ix["triangles"]
    assert ix["num_triangles"] == soup["num_triangles"] == len(T) > 1000
    assert T.min() >= 0 and T.max() < len(V)
    assert np.all((T[:, 0] != T[:, 1]) & (T[:, 1] != T[:, 2]) & (T[:, 0] != T[:, 2]))
    np.testing.assert_allclose(V[T].reshape(-1, 9), soup["tris"], rtol=0, atol=1e-5)
    e = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1).astype(np.int64)
    _, uses = np.unique(e[:, 0] * len(V) + e[:, 1], return_counts=True)
    assert uses.max() <= 2
    assert (uses == 2).mean() > 0.9
    assert len(np.unique(T)) > 0.9 * len(V)  # few unreferenced vertices



def test_mesh_over_capacity_reports_the_count_and_keeps_the_first_triangles():
    """Mesh buffers are sized in triangles (the reference's MeshEngine fills
    a fixed buffer): an extraction that does not fit returns YAS_E_CAPACITY
    with the full counts, and the triangles (and, indexed, the vertices)
    that fit are the first ones of the complete extraction, in order."""
    _, _, _, _, _, gm = fuse_both(FMTS[1])
    full = rf.MeshEngine(2_000_000)
    soup = full.Meshify(gm)
    ix = full.MeshifyIndexed(gm)
    n, nv = soup["num_triangles"], ix["num_vertices"]
    assert n > 3000
    cap = n // 3
    small = rf.MeshEngine(cap)
    with pytest.raises(native.YasError) as e:
        native.call("yas_mesh_extract", small.h, gm.h, ctypes.byref(ctypes.c_int()))
    assert e.value.code == 6  # YAS_E_CAPACITY
    part = small.Meshify(gm)
    assert part["num_triangles"] == n and len(part["tris"]) == cap
    assert_bits_equal(part["tris"], soup["tris"][:cap])
    assert_bits_equal(part["normals"], soup["normals"][:cap])
    pix = small.MeshifyIndexed(gm)
    assert pix["num_triangles"] == n and pix["num_vertices"] == nv
    kv = min(nv, 3 * cap)
    assert len(pix["triangles"]) == cap and len(pix["vertices"]) == kv
    np.testing.assert_array_equal(pix["triangles"], ix["triangles"][:cap])
    assert_bits_equal(pix["vertices"], ix["vertices"][:kv])

def test_pipeline_matches_oracle_pipeline():
    """The dense half of Tracking::TrackRGBD over 6 frames: GPU poses within
    1e-4 m / 1e-4 rad of the oracle's, same accept/reject decisions."""
    n = 6
    p, K, fr = frames(n)
    vox, mu = 0.008, 0.024
    ctx = rf.default_context()
    mp = rf.map_params(1 << 16, 1 << 15, vox, mu, rf.VOXEL_F32_F32)
    tp = rf.tracker_params(3, (10, 5, 3), rf.TRACK_RGBD, (1.0, 1.0, 100.0, 10.0))
    pl = rf.DensePipeline(p.width, p.height, K, mp, tp, ctx=ctx)
    op = po.OraclePipeline(p.width, p.height, K, vox, mu, 3, (10, 5, 3), po.MODE_RGBD)
    for t in range(n):
        d, rgb, _ = fr[t]
        pl.step(to_dev(d), to_dev(rgb))
        To, ok_o, _ = op.step(d, rgb)
        Tg, ok_g = pl.pose(t)
        assert ok_o == ok_g, t
        et, er = pose_err(Tg, To)
        assert et < 1e-4 and er < 1e-4, (t, et, er)


def test_pool_exhaustion_reports_and_keeps_a_subset():
    p, K, fr = frames(1)
    z = po.depth_u16_to_f32(fr[0][0])
    om = po.OracleMap(0.008, 0.024)
    om.fuse(z, K, np.eye(4))
    cap = 200
    gm = rf.MapStruct(K.astype(np.float32), 1 << 14, cap, 0.008, 0.024)
    gm.Fuse(to_dev(z), np.eye(4))
    st = gm.stats()
    assert st.num_allocated == cap
    assert st.flags & 1
    kg, _ = gm.export()
    ko = {tuple(k) for k in om.export(with_voxels=False)}
    assert {tuple(k) for k in kg} <= ko


def test_a_full_table_reports_the_probe_limit_and_keeps_a_subset():
    """A table of 64 slots fills before the frame's blocks are all inserted:
    the probe-limit flag is set (the pool has room), all 64 slots hold
    blocks, and each is a block the oracle allocates."""
    p, K, fr = frames(1)
    z = po.depth_u16_to_f32(fr[0][0])
    om = po.OracleMap(0.008, 0.024)
    om.fuse(z, K, np.eye(4))
    ko = {tuple(k) for k in om.export(with_voxels=False)}
    assert len(ko) > 64
    gm = rf.MapStruct(K.astype(np.float32), 64, 1 << 12, 0.008, 0.024)
    gm.Fuse(to_dev(z), np.eye(4))
    st = gm.stats()
    assert st.flags & 2 and not st.flags & 1, st.flags
    assert st.num_allocated == 64
    kg, _ = gm.export()
    assert {tuple(k) for k in kg} <= ko


def test_blocks_beyond_the_key_range_are_flagged_and_not_allocated():
    """Block coordinates are packed in 21 bits each (+-2^20 blocks); with the
    camera 100 km along x (1.56 M blocks of 6.4 cm) every block of the frame
    is out of range: the key-range flag is set and nothing is allocated."""
    p, K, fr = frames(1)
    z = po.depth_u16_to_f32(fr[0][0])
    gm = rf.MapStruct(K.astype(np.float32), 1 << 14, 1 << 12, 0.008, 0.024)
    T = np.eye(4)
    T[0, 3] = 1.0e5
    gm.Fuse(to_dev(z), T)
    st = gm.stats()
    assert st.flags & 4, st.flags
    assert st.num_allocated == 0


def test_empty_depth_frame_allocates_nothing_and_tracks_identity():
    p, K, fr = frames(1)
    W, H = p.width, p.height
    gm = rf.MapStruct(K.astype(np.float32), 1 << 14, 1 << 12, 0.008, 0.024)
    zeros = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    gm.Fuse(zeros, np.eye(4))
    assert gm.stats().num_allocated == 0
    trk = rf.RGBDTracking(W, H, K, 3, (10, 5, 3), rf.TRACK_DEPTH)
    trk.SetReferenceDepth(zeros)
    trk.SetReferenceModel(torch.zeros((H, W, 4), dtype=torch.float32, device="cuda"))
    trk.SetTrackingDepth(zeros)
    T, st = trk.GetTransform()
    assert st.good
    np.testing.assert_array_equal(T, np.eye(4))


FACADE_EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                          "facade_replay")


def test_cpp_facade_replays_track_rgbd():
    """The reference-shaped C++ facade on a wavy surface: tracks identity,
    fuses, merges a submap, meshes (built by __graft_entry__.build())."""
    assert os.path.exists(FACADE_EXE), "run __graft_entry__.build() first"
    r = subprocess.run([FACADE_EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tracked=1" in r.stdout


def test_cpp_facade_sequence_equals_the_pipeline(tmp_path):
    """The facade driven exactly as Tracking.cpp:139-197 drives RGBDFusion:
    host images (8-bit gray as System.cpp:47-49 makes it, float depth), the
    reference's default 5-level pyramid with {10,5,3,3,3} iterations and its
    1/1/100/10 weights, reference int16 + uint8 storage, the map pose through
    mTcw, rejected frames written out by WriteDebugImages.  Ten frames of the
    room at 320x240 with a 15 cm jump at frame 6.  Every pose and decision
    equals the C-ABI pipeline's on the same frames, bit for bit;
    GetRayTracingResultDepth equals the ray cast's z; GetReferenceDepth
    returns the depth just uploaded."""
    W, H, n = 320, 240, 10
    p = synth.params(width=W, height=H, theta0=np.pi / 4)
    K = synth.intrinsics(p)
    frames_ = [synth.frame(p, t)[:2] for t in range(6)]
    Tj = synth.pose(p, 5).copy()
    Tj[:3, 3] += 0.15 * Tj[:3, 0]
    frames_.append(synth.frame_at(p, Tj, 6))
    frames_ += [synth.frame(p, t)[:2] for t in range(6, 9)]
    path = tmp_path / "frames.bin"
    with open(path, "wb") as f:
        f.write(np.array([W, H, n], np.int32).tobytes())
        f.write(np.asarray(K, np.float64).tobytes())
        for d, c in frames_:
            f.write(po.rgb_to_intensity(c, mode=1).astype(np.uint8).tobytes())
            f.write(po.depth_u16_to_f32(d).tobytes())
    r = subprocess.run([FACADE_EXE, str(path), "5", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.split() for l in r.stdout.splitlines() if l.startswith("frame ")]
    info = dict(l.split() for l in r.stdout.splitlines() if not l.startswith("frame "))
    assert len(rows) == n
    mp = rf.map_params(1 << 18, 1 << 15, 0.01, 0.03, rf.VOXEL_I16_U8)
    tp = rf.tracker_params(5, (10, 5, 3, 3, 3), rf.TRACK_RGBD, (1.0, 1.0, 100.0, 10.0))
    pl = rf.DensePipeline(W, H, K, mp, tp, gray_mode=1)
    for d, c in frames_:
        pl.step(to_dev(d), to_dev(c))
    for t, row in enumerate(rows):
        Tf = np.array([float(v) for v in row[3:]]).reshape(4, 4)
        Tp, okp = pl.pose(t)
        assert int(row[2]) == int(okp), (t, row[2], okp)
        np.testing.assert_array_equal(Tf, Tp, err_msg="frame %d" % t)
    assert [int(r_[2]) for r_ in rows][6] == 0  # the jump
    for name in ("curr_image.pgm", "last_image.pgm", "gx.pgm", "gy.pgm"):
        assert (tmp_path / name).stat().st_size == len("P5\n%d %d\n255\n" % (W, H)) + W * H
    assert float(info["depth_max_abs_diff"]) == 0.0
    assert info["raw_depth_ok"] == "1" and int(info["triangles"]) > 10000
    pl.close()
    # the same sequence through the oracle's pipeline (float32): same
    # decisions (frames 6 and 7 rejected: after the jump the reference
    # intensity is the jump frame's), poses within 1e-4 m / 1e-4 rad
    op = po.OraclePipeline(W, H, K, 0.01, 0.03, 5, (10, 5, 3, 3, 3), po.MODE_RGBD, (1, 1, 100, 10),
                           voxfmt=po.VOX_REF, gray_mode=1)
    for t, (d, c) in enumerate(frames_):
        To, ok, _ = op.step(d, c)
        Tf = np.array([float(v) for v in rows[t][3:]]).reshape(4, 4)
        assert int(rows[t][2]) == int(ok), t
        et, er = pose_err(Tf, To)
        assert et <= 1e-4 and er <= 1e-4, (t, et, er)


def test_plain_method_sequence_against_oracle():
    """The plain method end to end: 30 frames at 640x480, voxel 8 mm,
    truncation 3 voxels, 2^20 buckets, f32 storage, 3-level pyramid (10/5/3)
    tracking depth and colour with the reference's weights (the mode of its
    shipped freiburg3 settings), against the float32 and the float64 oracle
    pipelines.  All three accept the same frames.  Each frame's tracked
    motion (T_{t-1}^-1 T_t) and every chained pose is within 1e-4 m / 1e-4
    rad of both oracles on every frame where the two agree to 1e-5.  Where
    they do not, float32 rounding alone decides an ICP association at the
    1 cm gate (on this sequence from frame 25 on, where the oracles part by
    1.4e-4 m): either side is the algorithm's answer, and the GPU, whose
    sums round in another order, must match one of the two.  The final
    volume shares its blocks with both oracles' and its observed voxels are
    within 0.1 voxel RMS of theirs (nearest-pixel depth lookups that round
    to another pixel in float32 give most of the difference to float64).

    Depth alone (ICP with the reference's fixed 1 cm association gate) does
    not track this trajectory: the oracle itself ends 2.5 cm from the truth
    after one frame and its float32 and float64 runs differ by millimetres,
    so poses of that mode are compared per step on identical inputs instead
    (the tracker tests)."""
    n, W, H = 30, 640, 480
    vox, mu = 0.008, 0.024
    p, K, fr = frames(n, W, H)
    mp = rf.map_params(1 << 20, 1 << 17, vox, mu, rf.VOXEL_F32_F32)
    tp = rf.tracker_params(3, (10, 5, 3), rf.TRACK_RGBD, (1.0, 1.0, 100.0, 10.0))
    pl = rf.DensePipeline(W, H, K, mp, tp)
    o32 = po.OraclePipeline(W, H, K, vox, mu, 3, (10, 5, 3), po.MODE_RGBD)
    o64 = po.OraclePipeline(W, H, K, vox, mu, 3, (10, 5, 3), po.MODE_RGBD, precision=64)
    poses = []
    for t in range(n):
        d, rgb, _ = fr[t]
        pl.step(to_dev(d), to_dev(rgb))
        T32, ok32, _ = o32.step(d, rgb)
        T64, ok64, _ = o64.step(d, rgb)
        Tg, okg = pl.pose(t)
        assert okg == ok32 == ok64, t
        poses.append((Tg, T32, T64))

    def motion(t, j):
        return np.linalg.inv(poses[t - 1][j]) @ poses[t][j]

    rows = []
    for t in range(1, n):
        g64 = max(pose_err(motion(t, 0), motion(t, 2)) + pose_err(poses[t][0], poses[t][2]))
        g32 = max(pose_err(motion(t, 0), motion(t, 1)) + pose_err(poses[t][0], poses[t][1]))
        o = max(pose_err(motion(t, 1), motion(t, 2)) + pose_err(poses[t][1], poses[t][2]))
        rows.append((t, g64, g32, o))
    print("plain method: frame, gpu-fp64, gpu-fp32 oracle, fp32-fp64 oracles:",
          " ".join("%d:%.1e/%.1e/%.1e" % r for r in rows))
    for t, g64, g32, o in rows:
        assert min(g64, g32) <= 1e-4, (t, g64, g32)
        if o <= 1e-5:
            assert g64 <= 1e-4 and g32 <= 1e-4, (t, g64, g32)
    kg, sg, wg = pl.export_map_sorted(rf.VOXEL_F32_F32)
    for name, op in (("fp32", o32), ("fp64", o64)):
        om = po.OracleMap.__new__(po.OracleMap)
        om.h = po.lib().orc_pipe_map(op.h)
        try:
            ko, so, wo = om.export()
        finally:
            om.h = None  # owned by the pipeline
        idx = {tuple(k): i for i, k in enumerate(ko)}
        common = [(i, idx[tuple(k)]) for i, k in enumerate(kg) if tuple(k) in idx]
        a = np.array([c[0] for c in common])
        b = np.array([c[1] for c in common])
        both = (wg[a] > 0) & (wo[b] > 0)
        rms = float(np.sqrt(np.mean((sg[a][both] - so[b][both]) ** 2))) * mu / vox
        print("plain method: blocks", len(kg), name, len(ko), "common", len(common), "tsdf rms (voxels)", rms)
        assert len(common) >= 0.999 * max(len(kg), len(ko)) and len(common) > 5000
        assert rms <= 0.1, (name, rms)


def test_depth_only_icp_sequence_against_oracle():
    """Depth alone, where the ICP term (RGBDTracking.cu:305-406) does all the
    work: 12 frames at 640x480 of noise-free depth, the camera moving 3 mm a
    frame, 8 mm voxels, truncation 3 voxels, 3 levels with 10/5/3 iterations.
    At the true motion two thirds of the valid level-0 pixels pass the 1 cm /
    20 degree association gates (against 7 % with the default sensor noise,
    whose per-pixel normals fail the angle gate), and the float64 oracle
    stays within 1 mm of the generator's poses.  Every frame's motion and
    chained pose on the GPU is within 1e-4 m / 1e-4 rad of the float32 and
    of the float64 oracle."""
    n, W, H = 12, 640, 480
    vox, mu = 0.008, 0.024
    p, K, fr = frames(n, W, H, step=0.003, noise_k=0.0)
    tr = po.OracleTracker(W, H, K, 1, (1,), po.MODE_DEPTH, precision=64)
    tr.set_reference_depth(po.depth_u16_to_f32(fr[0][0]))
    tr.set_tracking_depth(po.depth_u16_to_f32(fr[1][0]))
    true_dt = np.linalg.inv(np.linalg.inv(fr[0][2]) @ fr[1][2])
    sums, _ = tr.step_sums(0, true_dt, 0)
    valid = np.count_nonzero(fr[1][0])
    print("icp sequence: inliers at the true motion", int(sums[28]), "of", valid, "valid pixels")
    assert sums[28] >= 0.6 * valid

    mp = rf.map_params(1 << 20, 1 << 17, vox, mu, rf.VOXEL_F32_F32)
    tp = rf.tracker_params(3, (10, 5, 3), rf.TRACK_DEPTH)
    pl = rf.DensePipeline(W, H, K, mp, tp)
    o32 = po.OraclePipeline(W, H, K, vox, mu, 3, (10, 5, 3), po.MODE_DEPTH)
    o64 = po.OraclePipeline(W, H, K, vox, mu, 3, (10, 5, 3), po.MODE_DEPTH, precision=64)
    T0 = fr[0][2]
    poses = []
    for t in range(n):
        d, rgb, Tt = fr[t]
        pl.step(to_dev(d), to_dev(rgb))
        T32, ok32, _ = o32.step(d, rgb)
        T64, ok64, _ = o64.step(d, rgb)
        Tg, okg = pl.pose(t)
        assert okg and ok32 and ok64, t
        assert pose_err(T64, np.linalg.inv(T0) @ Tt)[0] <= 1e-3, t
        poses.append((Tg, T32, T64))
    worst = 0.0
    for t in range(1, n):
        for j in (1, 2):
            mg = np.linalg.inv(poses[t - 1][0]) @ poses[t][0]
            mo = np.linalg.inv(poses[t - 1][j]) @ poses[t][j]
            e = max(pose_err(mg, mo) + pose_err(poses[t][0], poses[t][j]))
            worst = max(worst, e)
            assert e <= 1e-4, (t, j, e)
    print("icp sequence: worst gpu-oracle pose error %.2e" % worst)
    pl.close()
