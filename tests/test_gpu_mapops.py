"""GPU tests of the volume-management entry points: map-to-map fusion
against the oracle (bit-exact block sets and voxels), growth, hibernation and
the on-disk format as exact round trips, and the PLY writer."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from yetanotherslam_amd import native
from yetanotherslam_amd import rgbdfusion as rf
from yetanotherslam_amd import synth

from helpers import assert_bits_equal, frames, to_dev

pytestmark = pytest.mark.gpu

FMTS = [(rf.VOXEL_I16_U8, po.VOX_REF), (rf.VOXEL_F32_F32, po.VOX_F32), (rf.VOXEL_F16_U16, po.VOX_F16)]
VOX, MU = 0.008, 0.024


def sorted_oracle(om):
    k, s, w = om.export()
    order = np.lexsort((k[:, 2], k[:, 1], k[:, 0]))
    return k[order], s[order].astype(np.float32), w[order].astype(np.float32)


def two_submaps(fmt_pair, n=4):
    """Submap A fuses frames 0-1 in the frame of camera 0, submap B frames
    2-3 in the frame of camera 2 (poses are camera -> world)."""
    gfmt, ofmt = fmt_pair
    p, K, fr = frames(n)
    gt = synth.relative_poses(p, n)
    Ta, Tb = gt[0], gt[2]
    maps = {}
    for name, T0, ts in (("a", Ta, (0, 1)), ("b", Tb, (2, 3))):
        om = po.OracleMap(VOX, MU, precision=32, voxfmt=ofmt)
        gm = rf.MapStruct(K.astype(np.float32), 1 << 16, 1 << 14, VOX, MU, voxel_format=gfmt)
        gm.SetPose(T0)
        for t in ts:
            z = po.depth_u16_to_f32(fr[t][0])
            Tl = np.linalg.inv(T0) @ gt[t]
            om.fuse(z, K, Tl)
            gm.Fuse(to_dev(z), Tl)
        maps[name] = (om, gm)
    torch.cuda.synchronize()
    return p, K, fr, gt, Ta, Tb, maps


@pytest.mark.parametrize("fmt", FMTS, ids=["i16u8", "f32", "f16u16"])
def test_map_merge_matches_oracle(fmt):
    """MapStruct::Fuse(MapStruct*) (VoxelMap.cu:400-449): the merged block
    set and every voxel are bit-exact against the oracle's merge."""
    _, _, _, _, Ta, Tb, maps = two_submaps(fmt)
    (oa, ga), (ob, gb) = maps["a"], maps["b"]
    nb_before = ga.stats().num_allocated
    created = oa.fuse_map(ob, Ta, Tb)
    ga.FuseMap(gb)
    keys, sdf, wt = ga.export_sorted()
    ko, so, wo = sorted_oracle(oa)
    assert created > 100 and len(keys) == nb_before + created
    np.testing.assert_array_equal(keys, ko)
    assert_bits_equal(sdf, so, "sdf")
    assert_bits_equal(wt, wo, "weight")
    # dst grew by the source's capacities first (VoxelMap.cu:412-417)
    P = rf.MapParams()
    native.call("yas_map_get_params", ga.h, rf.ctypes.byref(P), None)
    assert P.num_buckets == 2 * (1 << 16) and P.num_blocks == 2 * (1 << 14)


def test_reserve_hibernate_and_file_round_trips_are_exact(tmp_path):
    """Growth, hibernation and save/load keep every block and voxel; the
    rebuilt tables keep working (a further fuse matches the untouched map),
    and a hibernating map refuses volume calls."""
    gfmt = rf.VOXEL_F16_U16
    p, K, fr = frames(3)
    gt = synth.relative_poses(p, 3)
    z = [to_dev(po.depth_u16_to_f32(fr[t][0])) for t in range(3)]

    def build():
        m = rf.MapStruct(K.astype(np.float32), 1 << 15, 1 << 13, VOX, MU, voxel_format=gfmt)
        for t in range(2):
            m.Fuse(z[t], gt[t])
        return m

    ref, m = build(), build()
    k0, s0, w0 = ref.export_sorted()
    rc = rf.RayTraceEngine(p.width, p.height, K)
    rc.RayTrace(ref, gt[1])
    v0 = rc.GetVMap()

    m.Reserve(1 << 17, 1 << 15)
    P = rf.MapParams()
    native.call("yas_map_get_params", m.h, rf.ctypes.byref(P), None)
    assert (P.num_buckets, P.num_blocks) == (1 << 17, 1 << 15)
    for arr, ref_arr in zip(m.export_sorted(), (k0, s0, w0)):
        np.testing.assert_array_equal(arr, ref_arr)

    m.Hibernate()
    assert m.is_hibernating()
    with pytest.raises(native.YasError):
        m.Fuse(z[2], gt[2])
    m.ReActivate()
    assert not m.is_hibernating()
    for arr, ref_arr in zip(m.export_sorted(), (k0, s0, w0)):
        np.testing.assert_array_equal(arr, ref_arr)
    rc.RayTrace(m, gt[1])
    assert_bits_equal(rc.GetVMap(), v0, "vertex map after re-activation")

    m.SetPose(gt[1])
    path = tmp_path / "map.yas"
    m.SaveToFile(path)
    l = rf.MapStruct.ReadFromFile(path)
    assert l.voxel_format == gfmt
    np.testing.assert_array_equal(l.GetPose(), gt[1])
    for arr, ref_arr in zip(l.export_sorted(), (k0, s0, w0)):
        np.testing.assert_array_equal(arr, ref_arr)

    # the rebuilt tables take new blocks exactly like the original's
    for mm in (ref, m, l):
        mm.Fuse(z[2], gt[2])
    ke, se, we = ref.export_sorted()
    for mm in (m, l):
        k, s, w = mm.export_sorted()
        np.testing.assert_array_equal(k, ke)
        assert_bits_equal(s, se, "sdf after further fusion")
        assert_bits_equal(w, we, "weight after further fusion")


def test_ply_file_holds_the_indexed_mesh(tmp_path):
    p, K, fr = frames(2)
    gt = synth.relative_poses(p, 2)
    m = rf.MapStruct(K.astype(np.float32), 1 << 15, 1 << 13, VOX, MU)
    for t in range(2):
        m.Fuse(to_dev(po.depth_u16_to_f32(fr[t][0])), gt[t])
    me = rf.MeshEngine(1_000_000)
    ix = me.MeshifyIndexed(m)
    path = tmp_path / "mesh.ply"
    me.WritePLY(path)
    data = path.read_bytes()
    head_end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:head_end].decode()
    nv, nt = ix["num_vertices"], ix["num_triangles"]
    assert "element vertex %d" % nv in head and "element face %d" % nt in head and nt > 1000
    verts = np.frombuffer(data, np.float32, nv * 3, head_end).reshape(-1, 3)
    np.testing.assert_array_equal(verts, ix["vertices"])
    faces = np.frombuffer(data, np.uint8, nt * 13, head_end + nv * 12).reshape(nt, 13)
    assert np.all(faces[:, 0] == 3)
    np.testing.assert_array_equal(faces[:, 1:].copy().view(np.int32), ix["triangles"])
