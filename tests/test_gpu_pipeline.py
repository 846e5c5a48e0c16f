"""The per-frame dense step on the GPU against the oracle on the branches the
plain sequences never take: the |t| > 0.1 m rejection (Tracking.cpp:178-184)
and the NaN update (RGBDTracking.cu:187-191), in the single pipeline, the
tracker class and the serving engine.  Settings are the precision form's
(640x480, 5 mm, truncation 4 voxels, ICP + photometric with weight 0.1) at
the benchmark's table and pool sizes where the full resolution is used."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from yetanotherslam_amd import rgbdfusion as rf
from yetanotherslam_amd import synth

from helpers import assert_bits_equal, pose_err, to_dev

pytestmark = pytest.mark.gpu

W, H = 640, 480
VOX, MU = 0.005, 0.02
ITERS = (10, 5, 3)
WEIGHTS = (0.1, 0.1, 1.0, 1.0)


def test_pipeline_rejects_a_jump_like_the_oracle():
    """Frames 0-3 of the room, then the room seen from 15 cm to the side of
    frame 3, then frames 4 and 5.  The jump is rejected on both sides: the
    pose stays, nothing is allocated or fused.  Every decision matches the
    float32 oracle's and every pose is within 1e-4 m / 1e-4 rad of it.
    2^22 buckets and 2^19 blocks, as the benchmark."""
    p = synth.params(W, H)
    K = synth.intrinsics(p)
    seq = [synth.frame(p, t)[:2] for t in range(4)]
    Tj = synth.pose(p, 3).copy()
    Tj[:3, 3] += 0.15 * Tj[:3, 0]
    seq.append(synth.frame_at(p, Tj, 4))
    seq += [synth.frame(p, t)[:2] for t in (4, 5)]
    mp = rf.map_params(1 << 22, 1 << 19, VOX, MU, rf.VOXEL_F32_F32)
    tp = rf.tracker_params(3, ITERS, rf.TRACK_RGBD, WEIGHTS)
    pl = rf.DensePipeline(W, H, K, mp, tp)
    op = po.OraclePipeline(W, H, K, VOX, MU, 3, ITERS, po.MODE_RGBD, WEIGHTS)
    decisions = []
    for t, (d, c) in enumerate(seq):
        if t == 4:
            blocks_before = pl.map_stats().num_allocated
        pl.step(to_dev(d), to_dev(c))
        To, ok_o, _ = op.step(d, c)
        Tg, ok_g = pl.pose(t)
        decisions.append((ok_g, ok_o))
        assert ok_g == ok_o, (t, ok_g, ok_o)
        et, er = pose_err(Tg, To)
        assert et <= 1e-4 and er <= 1e-4, (t, et, er)
        if t == 4:
            assert pl.map_stats().num_allocated == blocks_before
            np.testing.assert_array_equal(Tg, pl.pose(3)[0])
    print("jump sequence decisions (gpu, oracle):", decisions)
    assert decisions[4] == (False, False)
    pl.close()


def test_tracker_nan_update_returns_identity_and_keeps_its_buffers():
    """RGBDTracking.cu:187-191 through the tracker class: a current image
    scaled by 1e30 overflows the photometric J^T J; GetTransform returns the
    identity with IsTrackingGood() false and does not swap the buffers.  The
    same frames at scale 1 track and swap."""
    w, h = 160, 120
    p = synth.params(w, h)
    K = synth.intrinsics(p)
    (d0, c0, _), (d1, c1, _) = synth.frame(p, 0), synth.frame(p, 1)
    g0, g1 = po.rgb_to_intensity(c0), po.rgb_to_intensity(c1)
    z0, z1 = po.depth_u16_to_f32(d0), po.depth_u16_to_f32(d1)
    trk = rf.RGBDTracking(w, h, K, 3, ITERS, rf.TRACK_RGB)
    trk.SetReferenceImage(to_dev(g0))
    trk.SetReferenceDepth(to_dev(z0))
    trk.SetTrackingImage(to_dev(g1 * np.float32(1e30)))
    trk.SetTrackingDepth(to_dev(z1))
    ref_before, cur_before = trk.level(1, 0), trk.level(0, 0)
    T, st = trk.GetTransform(swap=True)
    assert not st.good and not trk.IsTrackingGood()
    np.testing.assert_array_equal(T, np.eye(4))
    assert_bits_equal(trk.level(1, 0), ref_before, "reference intensity kept")
    assert_bits_equal(trk.level(0, 0), cur_before, "current intensity kept")
    trk.SetTrackingImage(to_dev(g1))
    T, st = trk.GetTransform(swap=True)
    assert st.good and trk.IsTrackingGood()
    assert_bits_equal(trk.level(1, 0), g1, "swapped")
    trk.close()


def nan_sequence(n, w=160, h=120):
    p = synth.params(w, h)
    return p, synth.intrinsics(p), [synth.frame(p, t)[:2] for t in range(n)]


def oracle_identity_map(K, seq, vox, mu):
    om = po.OracleMap(vox, mu)
    for d, _ in seq:
        om.fuse(po.depth_u16_to_f32(d), K, np.eye(4))
    return om.export()


def test_pipeline_fuses_at_the_last_pose_after_nan_updates():
    """A photometric weight of 3e38 makes every solve NaN (w * H_rgb
    overflows in float32).  TrackRGBD accepts the identity each time
    (Tracking.cpp:175-194): every pose stays the identity, every frame is
    fused there (the volume equals the oracle's fused at the identity, bit for
    bit) and the reference intensity stays frame 0's, as in the oracle."""
    n, vox, mu = 4, 0.01, 0.03
    p, K, seq = nan_sequence(n)
    mp = rf.map_params(1 << 16, 1 << 14, vox, mu, rf.VOXEL_F32_F32)
    tp = rf.tracker_params(3, ITERS, rf.TRACK_RGBD, (3e38, 3e38, 1.0, 1.0))
    pl = rf.DensePipeline(p.width, p.height, K, mp, tp)
    op = po.OraclePipeline(p.width, p.height, K, vox, mu, 3, ITERS, po.MODE_RGBD, (3e38, 3e38, 1.0, 1.0))
    for t, (d, c) in enumerate(seq):
        pl.step(to_dev(d), to_dev(c))
        To, ok_o, _ = op.step(d, c)
        Tg, ok_g = pl.pose(t)
        assert ok_g and ok_o, t
        np.testing.assert_array_equal(Tg, np.eye(4))
        np.testing.assert_array_equal(To, np.eye(4))
    kg, sg, wg = pl.export_map_sorted(rf.VOXEL_F32_F32)
    ko, so, wo = oracle_identity_map(K, seq, vox, mu)
    np.testing.assert_array_equal(kg, ko)
    assert_bits_equal(sg, so, "tsdf")
    assert_bits_equal(wg, wo, "weight")
    assert_bits_equal(pl.tracker_level(1, 0), po.rgb_to_intensity(seq[0][1]), "reference intensity")
    pl.close()


def test_server_fuses_at_the_last_pose_after_nan_updates():
    """The same NaN sequence through the serving engine's batched kernels:
    identity poses, every frame fused, the volume equal to the oracle's."""
    n, vox, mu = 4, 0.01, 0.03
    p, K, seq = nan_sequence(n)
    tp = rf.tracker_params(3, ITERS, rf.TRACK_RGBD, (3e38, 3e38, 1.0, 1.0))
    sv = rf.Server(p.width, p.height, 4, tp, rf.VOXEL_F32_F32, hbm_budget=8 << 30)
    sid = sv.add_stream(p.width, p.height, K, rf.map_params(1 << 16, 1 << 14, vox, mu, rf.VOXEL_F32_F32))
    dev = [(to_dev(d), to_dev(c)) for d, c in seq]
    torch.cuda.synchronize()
    for d, c in dev:
        sv.submit(sid, d, c)
    sv.drain()
    res = sorted(sv.poll(), key=lambda r: r["request"])
    assert len(res) == n
    for r in res:
        assert r["tracked"]
        np.testing.assert_array_equal(r["pose"], np.eye(4))
    kg, sg, wg = sv.stream_map(sid).export_sorted()
    ko, so, wo = oracle_identity_map(K, seq, vox, mu)
    np.testing.assert_array_equal(kg, ko)
    assert_bits_equal(sg, so, "tsdf")
    assert_bits_equal(wg, wo, "weight")
    sv.close()


def test_pipeline_timing_samples_stages_and_kernels():
    """yas_pipeline_enable_timing(n): every timed frame's start and end, and
    every n-th frame's stage boundaries and single-kernel durations."""
    import ctypes
    from yetanotherslam_amd import native
    n, w, h = 7, 160, 120
    p, K, seq = nan_sequence(n, w, h)
    pl = rf.DensePipeline(w, h, K, rf.map_params(1 << 16, 1 << 14, 0.01, 0.03), rf.tracker_params())
    pl.step(*(to_dev(a) for a in seq[0]))
    native.call("yas_pipeline_enable_timing", pl.h, 2)
    for d, c in seq[1:]:
        pl.step(to_dev(d), to_dev(c))
    stage = np.zeros(8)
    frame = np.zeros(8)
    nt = ctypes.c_int()
    native.call("yas_pipeline_timing", pl.h, stage.ctypes.data, frame.ctypes.data, 8, ctypes.byref(nt))
    kern = np.zeros(4)
    ns = ctypes.c_int()
    native.call("yas_pipeline_kernel_timing", pl.h, kern.ctypes.data, ctypes.byref(ns))
    assert nt.value == n - 1 and ns.value == 3  # frames 1..6 timed, 1, 3, 5 sampled
    assert np.all(frame[:n - 1] > 0) and np.all(stage > 0) and np.all(kern > 0)
    # each kernel runs inside its stage
    for k, i in ((0, 2), (1, 5), (2, 6), (3, 7)):
        assert kern[k] <= stage[i] * 1.05 + 1e-3, (k, kern[k], stage[i])
    pl.close()
