"""The benchmark's own configurations under test (yetanotherslam_amd/
workloads.py defines them for both):

* marching cubes over the room fused from 300 frames at 5 mm (about 75 k
  blocks, 5 M triangles) in float32 and float16 sdf: block keys, cube cases,
  triangle counts, vertices and normals bit-exact against the oracle, which
  fuses the same frames itself (so the fusion of the whole volume is checked
  too);
* the batching form's 64 submaps (5-60 frames, 640x480 and 320x240,
  2^18 buckets each): the captured batch leaves every submap bit-identical to
  fusing it one call at a time, and the smallest submap of each resolution
  is bit-identical to the oracle's;
* the serving form's 32 streams at 640x480 (float32 and float16 storage) and
  in the seeded 640x480 / 320x240 mix, submitted a frame of every stream per
  round so each batch holds 32 frames: every stream's poses, decisions and
  volume are bit-identical to its own single pipeline, and one 640x480
  stream's poses are within 1e-4 m / 1e-4 rad of the oracle's pipeline with
  the same decisions.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from yetanotherslam_amd import rgbdfusion as rf
from yetanotherslam_amd import synth
from yetanotherslam_amd import workloads as wl

from helpers import assert_bits_equal, pose_err, to_dev

pytestmark = pytest.mark.gpu

MESH_FRAMES = 300
NSEQ = 301  # frames of the benchmark's default run, from which segments are drawn


class FrameCache:
    """Synthetic frames of the room sequence at one resolution, made once."""

    def __init__(self, width, height):
        self.p = wl.synth_params(width, height)
        self.K = synth.intrinsics(self.p)
        self.host, self.dev = {}, {}

    def get(self, t):
        if t not in self.host:
            d, c, _ = synth.frame(self.p, t)
            self.host[t] = (d, c)
            self.dev[t] = (to_dev(d), to_dev(c))
        return self.host[t], self.dev[t]


@pytest.fixture(scope="module")
def hi_frames():
    return FrameCache(wl.W, wl.H)


@pytest.fixture(scope="module")
def lo_frames():
    return FrameCache(wl.W // 2, wl.H // 2)


# ------------------------------------------------------------------ marching cubes

@pytest.fixture(scope="module")
def fused_room(hi_frames):
    """The room fused from 300 frames at the true poses, by the GPU (float32
    and float16 sdf) and by the oracle (both storages, one thread each)."""
    p, K = hi_frames.p, hi_frames.K
    gt = synth.relative_poses(p, MESH_FRAMES)
    depth = [po.depth_u16_to_f32(synth.frame(p, t)[0]) for t in range(MESH_FRAMES)]
    fmts = {"f32": (rf.VOXEL_F32_F32, po.VOX_F32), "f16u16": (rf.VOXEL_F16_U16, po.VOX_F16)}

    def oracle(ofmt):
        om = po.OracleMap(wl.VOXEL, wl.TRUNC, precision=32, voxfmt=ofmt)
        for t in range(MESH_FRAMES):
            om.fuse(depth[t], K, gt[t])
        return om

    with ThreadPoolExecutor(2) as ex:  # the oracle's calls release the GIL
        futs = {k: ex.submit(oracle, v[1]) for k, v in fmts.items()}
        gpu = {}
        for k, (gfmt, _) in fmts.items():
            m = rf.MapStruct(K.astype(np.float32), wl.BUCKETS, wl.BLOCKS, wl.VOXEL, wl.TRUNC, voxel_format=gfmt)
            for t in range(MESH_FRAMES):
                m.Fuse(to_dev(depth[t]), gt[t])
            gpu[k] = m
        torch.cuda.synchronize()
        orc = {k: f.result() for k, f in futs.items()}
    yield gpu, orc
    for m in gpu.values():
        m.close()


@pytest.mark.parametrize("fmt", ["f32", "f16u16"])
def test_mesh_of_the_300_frame_room_bit_exact(fused_room, fmt):
    gpu, orc = fused_room
    gm, om = gpu[fmt], orc[fmt]
    mo = om.mesh()
    me = rf.MeshEngine(8_000_000)
    mg = me.Meshify(gm, with_cases=True)
    keys, counts, cases, tris, normals = rf.canonical_mesh(mg)
    print("300-frame room, %s: %d blocks, %d triangles" % (fmt, len(keys), mg["num_triangles"]))
    assert len(keys) > 70000 and mg["num_triangles"] > 4_000_000
    np.testing.assert_array_equal(mo["keys"], keys)
    np.testing.assert_array_equal(mo["cases"], cases)
    np.testing.assert_array_equal(mo["counts"], counts)
    assert mg["num_triangles"] == len(mo["tris"])
    assert_bits_equal(mo["tris"], tris, "vertices")
    assert_bits_equal(mo["normals"], normals, "normals")
    me.close()


def test_mesh_of_a_c3_sized_volume_bit_exact():
    """Marching cubes at the size c3 names (200k-400k blocks at 5 mm): the
    wavy wall of workloads.wall_frames (a 7 x 6 grid of camera positions
    2.4 m x 1.8 m apart), in float16 sdf + uint16 weight storage.  The oracle meshes the same volume
    (imported from the device): the block set, every case index and
    triangle count, and every vertex and normal are bit-exact."""
    K = synth.intrinsics(wl.synth_params())
    m = rf.MapStruct(K.astype(np.float32), wl.BUCKETS, wl.BLOCKS, wl.VOXEL, wl.TRUNC, voxel_format=rf.VOXEL_F16_U16)
    for d, T in wl.wall_frames(K):
        m.Fuse(to_dev(d), T)
    torch.cuda.synchronize()
    keys, sdf, wt = m.export_sorted()
    assert len(keys) >= 200_000, len(keys)
    om = po.OracleMap(wl.VOXEL, wl.TRUNC, precision=32, voxfmt=po.VOX_F16)
    om.import_blocks(keys, sdf.astype(np.float64), wt.astype(np.float64))
    del sdf, wt
    mo = om.mesh()
    me = rf.MeshEngine(len(mo["tris"]) + 1)
    mg = me.Meshify(m, with_cases=True)
    gk, counts, cases, tris, normals = rf.canonical_mesh(mg)
    print("c3-sized volume: %d blocks, %d triangles" % (len(keys), mg["num_triangles"]))
    np.testing.assert_array_equal(mo["keys"], gk)
    np.testing.assert_array_equal(mo["cases"], cases)
    np.testing.assert_array_equal(mo["counts"], counts)
    assert mg["num_triangles"] == len(mo["tris"])
    assert_bits_equal(mo["tris"], tris, "vertices")
    assert_bits_equal(mo["normals"], normals, "normals")
    me.close()
    m.close()


# ------------------------------------------------------------------ batching

def test_64_submaps_batched_equal_one_at_a_time_and_the_oracle(hi_frames, lo_frames):
    sch = wl.BatchSchedule(NSEQ)
    poses = sch.poses()
    cache = {True: hi_frames, False: lo_frames}
    depth = [[None if sch.frame_index(s, i) is None else cache[bool(sch.hi[i])].get(sch.frame_index(s, i))[1][0]
              for i in range(sch.n)] for s in range(sch.steps)]
    torch.cuda.synchronize()

    def new_maps():
        return [rf.MapStruct(sch.intrinsics(i).astype(np.float32), wl.BATCH_BUCKETS, sch.map_blocks(i), wl.VOXEL,
                             wl.TRUNC) for i in range(sch.n)]

    seq = new_maps()
    for s in range(sch.steps):
        for i in range(sch.n):
            if depth[s][i] is not None:
                seq[i].FuseU16(depth[s][i], poses[s][i])
    bat = new_maps()
    b = rf.FusionBatch(bat)
    b.plan(depth, poses)
    cfg = b.autotune(reset_maps=True)
    b.run()
    torch.cuda.synchronize()
    print("64 submaps: %d frames (%d at 640x480), autotuned %s" % (int(sch.nfr.sum()), int(sch.nfr[sch.hi].sum()), cfg))
    total = 0
    for i in range(sch.n):
        ks, ss, ws = seq[i].export_sorted()
        kb, sb, wb = bat[i].export_sorted()
        np.testing.assert_array_equal(kb, ks, err_msg="map %d" % i)
        assert_bits_equal(sb, ss, "map %d sdf" % i)
        assert_bits_equal(wb, ws, "map %d weight" % i)
        total += len(kb)
    assert total > 500000
    for res in (True, False):
        i = int(np.flatnonzero(sch.hi == res)[np.argmin(sch.nfr[sch.hi == res])])
        om = po.OracleMap(wl.VOXEL, wl.TRUNC)
        for s in range(int(sch.nfr[i])):
            d = cache[res].get(sch.frame_index(s, i))[0][0]
            om.fuse(po.depth_u16_to_f32(d), sch.intrinsics(i), poses[s][i])
        ko, so, wo = om.export()
        kb, sb, wb = bat[i].export_sorted()
        np.testing.assert_array_equal(kb, ko, err_msg="map %d vs oracle" % i)
        assert_bits_equal(sb, so, "map %d sdf vs oracle" % i)
        assert_bits_equal(wb, wo, "map %d weight vs oracle" % i)
    b.close()
    for m in seq + bat:
        m.close()


# ------------------------------------------------------------------ serving

SERVE_FRAMES = 6


@pytest.mark.parametrize("mix,fmt", [("640x480", rf.VOXEL_F32_F32), ("640x480", rf.VOXEL_F16_U16),
                                     ("mixed", rf.VOXEL_F32_F32)], ids=["640x480-f32", "640x480-f16u16", "mixed-f32"])
def test_32_streams_each_equal_their_own_pipeline(hi_frames, lo_frames, mix, fmt):
    rng, hi = wl.serving_resolutions()
    if mix == "640x480":
        hi[:] = True
    start = wl.serving_starts(rng, NSEQ, SERVE_FRAMES)
    cache = {True: hi_frames, False: lo_frames}

    def frame(i, k):
        return cache[bool(hi[i])].get(int(start[i]) + k)

    for i in range(wl.SERVE_STREAMS):
        for k in range(SERVE_FRAMES):
            frame(i, k)
    torch.cuda.synchronize()
    mp = rf.map_params(wl.SERVE_BUCKETS, wl.SERVE_BLOCKS, wl.VOXEL, wl.TRUNC, fmt)
    tp = wl.tracker_params()
    sv = rf.Server(wl.W, wl.H, wl.SERVE_STREAMS, tp, voxel_format=fmt, hbm_budget=200 << 30)
    ids = [sv.add_stream(cache[bool(hi[i])].p.width, cache[bool(hi[i])].p.height, cache[bool(hi[i])].K, mp)
           for i in range(wl.SERVE_STREAMS)]
    sv.prepare()
    for k in range(SERVE_FRAMES):
        for i, sid in enumerate(ids):
            sv.submit(sid, *frame(i, k)[1])
        sv.pump()
    sv.drain()
    res = {sid: [] for sid in ids}
    for r in sv.poll(wl.SERVE_STREAMS * SERVE_FRAMES + 1):
        res[r["stream"]].append(r)
    st = sv.stats()
    print("32 streams %s fmt %d: %d frames in %d batches" % (mix, fmt, st["frames_done"], st["batches_done"]))
    assert st["frames_done"] == wl.SERVE_STREAMS * SERVE_FRAMES
    assert st["batches_done"] <= SERVE_FRAMES + 2  # whole rounds of 32 frames per batch
    for i, sid in enumerate(ids):
        c = cache[bool(hi[i])]
        pl = rf.DensePipeline(c.p.width, c.p.height, c.K, mp, tp)
        for k in range(SERVE_FRAMES):
            pl.step(*frame(i, k)[1])
        rs = sorted(res[sid], key=lambda r: r["request"])
        assert len(rs) == SERVE_FRAMES
        for k in range(SERVE_FRAMES):
            T, ok = pl.pose(k)
            assert rs[k]["tracked"] == ok, (i, k)
            np.testing.assert_array_equal(rs[k]["pose"], T, err_msg="stream %d frame %d" % (i, k))
        for a, b in zip(sv.stream_map(sid, fmt).export_sorted(), pl.export_map_sorted(fmt)):
            np.testing.assert_array_equal(a, b, err_msg="stream %d volume" % i)
        pl.close()
    # one 640x480 stream against the oracle's pipeline
    i = int(np.flatnonzero(hi)[0])
    ofmt = {rf.VOXEL_F32_F32: po.VOX_F32, rf.VOXEL_F16_U16: po.VOX_F16}[fmt]
    op = po.OraclePipeline(wl.W, wl.H, hi_frames.K, wl.VOXEL, wl.TRUNC, wl.LEVELS, wl.ITERS, po.MODE_RGBD,
                           wl.WEIGHTS, voxfmt=ofmt)
    rs = sorted(res[ids[i]], key=lambda r: r["request"])
    for k in range(SERVE_FRAMES):
        To, ok = op.step(*frame(i, k)[0])[:2]
        assert rs[k]["tracked"] == ok, k
        et, er = pose_err(rs[k]["pose"], To)
        assert et <= 1e-4 and er <= 1e-4, (k, et, er)
    sv.close()
