"""Pins of the oracle's per-frame dense step (CPU only): the pipeline that
replays FullSystem/src/Tracking.cpp:139-197 on one volume, checked against
the synthetic generator's ground truth; its |t| > 0.1 m rejection gate; the
NaN-update branch of RGBDTracking.cu:187-191; and OpenCV's 8-bit gray
conversion of System.cpp:47-49.

These check the oracle against what the mathematics and the reference's
control flow fix, not against the GPU: the GPU pipeline is compared with this
oracle by the GPU tests, so a composition slip made on both sides would pass
those and must fail here."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yetanotherslam_amd import synth

from helpers import pose_err

# The precision-form settings: 640x480, 5 mm voxels, truncation 4 voxels,
# 3 levels of 10/5/3 iterations, ICP + photometric with photometric weight 0.1.
W, H = 640, 480
VOX, MU = 0.005, 0.02
ITERS = (10, 5, 3)
WEIGHTS = (0.1, 0.1, 1.0, 1.0)
ROLL = 0.03  # rad per frame about the optical axis


def rolling_sequence(n):
    """Frames of the room seen from the circular path with a pitch wobble,
    a lift wobble and a steady roll, and ground truth in the frame of camera 0.
    On the plain circle every frame-to-frame motion is a power of one screw
    motion, so T_last * D and D * T_last coincide; the roll makes them differ."""
    p = synth.params(W, H)
    K = synth.intrinsics(p)
    T0inv = np.linalg.inv(synth.wobble_pose(p, 0, roll=ROLL))
    out = []
    for t in range(n):
        Tw = synth.wobble_pose(p, t, roll=ROLL)
        d, c = synth.frame_at(p, Tw, t)
        out.append((d, c, T0inv @ Tw))
    return p, K, out


def test_pipeline_follows_ground_truth_and_composes_T_last_times_DT_inverse():
    """FP64 oracle over 31 frames: every frame is accepted and its tracked
    motion T_{t-1}^-1 T_t is within 5 mm / 2 mrad of the true motion.

    Measured on this sequence: the largest motion error is 2.5 mm / 0.8 mrad.
    The two composition slips give (from the same tracker outputs):
    T_last * DT instead of T_last * DT^-1 is 35-46 mm off on every frame,
    and DT^-1 * T_last is 9.7-16 mm off on every frame from 24 on, so each
    fails this bound (Tracking.cpp:186 fixes T_w_cur = T_w_last * DT^-1)."""
    n = 31
    p, K, fr = rolling_sequence(n)
    pl = po.OraclePipeline(W, H, K, VOX, MU, 3, ITERS, po.MODE_RGBD, weights=WEIGHTS, precision=64)
    prev = None
    worst = (0.0, 0.0)
    for t, (d, c, Tgt) in enumerate(fr):
        T, ok, _ = pl.step(d, c)
        assert ok, t
        if t == 0:
            np.testing.assert_array_equal(T, np.eye(4))
        else:
            et, er = pose_err(np.linalg.inv(prev[0]) @ T, np.linalg.inv(prev[1]) @ Tgt)
            worst = (max(worst[0], et), max(worst[1], er))
            assert et <= 5e-3 and er <= 2e-3, (t, et, er)
        prev = (T, Tgt)
    print("oracle pipeline vs ground truth, worst per-frame motion error: %.2e m, %.2e rad" % worst)


def test_pipeline_rejects_a_jump_larger_than_10_cm_and_does_not_fuse_it():
    """Tracking.cpp:178-184: a frame whose tracked translation exceeds 0.1 m
    is rejected: the pose stays at the last frame's and nothing is fused.
    The jump frame is the room seen from 15 cm to the side of frame 3; the
    tracker resolves that translation, so the gate fires."""
    p = synth.params(W, H)
    K = synth.intrinsics(p)
    pl = po.OraclePipeline(W, H, K, VOX, MU, 3, ITERS, po.MODE_RGBD, weights=WEIGHTS)
    om = po.OracleMap.__new__(po.OracleMap)
    om.h = po.lib().orc_pipe_map(pl.h)
    try:
        for t in range(4):
            d, c, _ = synth.frame(p, t)
            T3, ok, _ = pl.step(d, c)
            assert ok, t
        before = om.export()
        Tj = synth.pose(p, 3).copy()
        Tj[:3, 3] += 0.15 * Tj[:3, 0]
        d, c = synth.frame_at(p, Tj, 4)
        Tr, ok, _ = pl.step(d, c)
        assert not ok
        np.testing.assert_array_equal(Tr, T3)
        after = om.export()
        for a, b in zip(before, after):
            np.testing.assert_array_equal(a, b)
    finally:
        om.h = None  # owned by the pipeline


def test_nan_update_returns_identity_and_keeps_the_reference_buffers():
    """RGBDTracking.cu:187-191: on a NaN update GetTransform returns the
    identity, sets the tracking-good flag false and returns before
    SwapFrameBuffer.  A current image scaled by 1e30 makes the photometric
    J^T J overflow to infinity in float32, and the LDLT solve of an infinite
    system yields NaN."""
    p = synth.params(160, 120)
    K = synth.intrinsics(p)
    d0, c0, _ = synth.frame(p, 0)
    d1, c1, _ = synth.frame(p, 1)
    g0, g1 = po.rgb_to_intensity(c0), po.rgb_to_intensity(c1)
    z0, z1 = po.depth_u16_to_f32(d0), po.depth_u16_to_f32(d1)
    trk = po.OracleTracker(160, 120, K, 3, ITERS, po.MODE_RGB)
    trk.set_reference_image(g0)
    trk.set_reference_depth(z0)
    trk.set_tracking_image(g1 * np.float32(1e30))
    trk.set_tracking_depth(z1)
    ref_int, cur_int = trk.level(1, 0).copy(), trk.level(0, 0).copy()
    T, good = trk.get_transform(swap=True)
    assert not good
    np.testing.assert_array_equal(T, np.eye(4))
    np.testing.assert_array_equal(trk.level(1, 0), ref_int)  # not swapped
    np.testing.assert_array_equal(trk.level(0, 0), cur_int)
    # the same inputs at 1x converge, and the swap happens
    trk.set_tracking_image(g1)
    T, good = trk.get_transform(swap=True)
    assert good and pose_err(T, np.eye(4))[0] > 1e-3
    np.testing.assert_array_equal(trk.level(1, 0), g1)


def test_pipeline_fuses_at_the_last_pose_after_a_nan_update():
    """Tracking.cpp:175-194 gates only on |t|: the identity that a NaN update
    returns is accepted, T_w_cur = T_w_last * I^-1 = T_w_last, and the frame is
    fused there.  A photometric weight of 3e38 makes every solve NaN in
    float32 (w * H_rgb overflows); the volume must then equal one fused at the
    identity for every frame, and the reference intensity stays frame 0's."""
    n, w, h = 4, 160, 120
    p = synth.params(w, h)
    K = synth.intrinsics(p)
    pl = po.OraclePipeline(w, h, K, 0.01, 0.03, 3, ITERS, po.MODE_RGBD, weights=(3e38, 3e38, 1.0, 1.0))
    om = po.OracleMap(0.01, 0.03)
    for t in range(n):
        d, c, _ = synth.frame(p, t)
        T, ok, _ = pl.step(d, c)
        assert ok, t
        np.testing.assert_array_equal(T, np.eye(4))
        om.fuse(po.depth_u16_to_f32(d), K, np.eye(4))
    view = po.OracleMap.__new__(po.OracleMap)
    view.h = po.lib().orc_pipe_map(pl.h)
    trk = po.OracleTracker.__new__(po.OracleTracker)
    trk.h, trk.W, trk.H, trk.levels = po.lib().orc_pipe_tracker(pl.h), w, h, 3
    try:
        for a, b in zip(view.export(), om.export()):
            np.testing.assert_array_equal(a, b)
        g0 = po.rgb_to_intensity(synth.frame(p, 0)[1])
        np.testing.assert_array_equal(trk.level(1, 0), g0)
    finally:
        view.h = None  # owned by the pipeline
        trk.h = None


def test_gray_mode_1_is_opencvs_14_bit_fixed_point_luma():
    """System.cpp:47-49 converts colour with cv::cvtColor(RGB2GRAY) on 8-bit
    images.  OpenCV documents Y = 0.299 R + 0.587 G + 0.114 B and computes it
    in 14-bit fixed point with round-half-up: (R*4899 + G*9617 + B*1868 +
    2^13) >> 14, whose weights sum to 2^14.  Known answers from the
    documented weights (each primary at 255 gives round(255 * weight)),
    channel order (a BGR slip swaps 76 and 29), and the whole 2^24 input cube
    against the real-valued luma: never more than 0.5 + 2^-14 * 255 away."""
    kat = {(0, 0, 0): 0, (255, 255, 255): 255, (255, 0, 0): 76, (0, 255, 0): 150, (0, 0, 255): 29,
           (128, 128, 128): 128, (1, 1, 1): 1, (10, 200, 30): 124, (200, 10, 30): 69}
    rgb = np.array(list(kat), np.uint8).reshape(1, -1, 3)
    got = po.rgb_to_intensity(rgb, mode=1)[0]
    np.testing.assert_array_equal(got, np.array(list(kat.values()), np.float32))
    # exhaustive: 256 slices of 256 x 256
    r = np.arange(256, dtype=np.uint8)
    gg, bb = np.meshgrid(r, r, indexing="ij")
    worst = 0.0
    for R in range(256):
        img = np.stack([np.full_like(gg, R), gg, bb], -1)
        y = po.rgb_to_intensity(img, mode=1)
        exact = 0.299 * R + 0.587 * gg.astype(np.float64) + 0.114 * bb.astype(np.float64)
        worst = max(worst, float(np.abs(y - exact).max()))
        assert np.all(y == np.floor(y)) and y.min() >= 0 and y.max() <= 255
    assert worst <= 0.5 + 255.0 / 16384, worst
