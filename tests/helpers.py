"""Shared helpers of the test suite: synthetic frames, host/device transfers
and exact comparisons."""
import numpy as np

from yetanotherslam_amd import synth

HALF_PI = np.pi / 2


def frames(n, width=320, height=240, theta0=np.pi / 4, seed=0, **kw):
    p = synth.params(width=width, height=height, theta0=theta0, seed=seed, **kw)
    K = synth.intrinsics(p)
    out = [synth.frame(p, t) for t in range(n)]
    return p, K, out


def to_dev(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        a = a.view(np.int16)
    return torch.from_numpy(a).cuda()


def bits32(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def assert_bits_equal(a, b, what=""):
    """Bit-exact float32 equality (NaN patterns aside)."""
    a32, b32 = np.asarray(a, np.float32), np.asarray(b, np.float32)
    assert a32.shape == b32.shape, (what, a32.shape, b32.shape)
    both_nan = np.isnan(a32) & np.isnan(b32)
    diff = (bits32(a32) != bits32(b32)) & ~both_nan
    if diff.any():
        i = np.argwhere(diff)[:5]
        raise AssertionError("%s: %d of %d values differ, first at %s: %s vs %s" % (
            what, int(diff.sum()), diff.size, i.tolist(), a32[tuple(i[0])], b32[tuple(i[0])]))


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-12, float(np.max(np.abs(b)))))


def pose_err(A, B):
    """(translation error m, rotation error rad) between two 4x4 poses."""
    E = np.linalg.inv(B) @ A
    c = np.clip((np.trace(E[:3, :3]) - 1.0) / 2.0, -1.0, 1.0)
    return float(np.linalg.norm(E[:3, 3])), float(np.arccos(c))


def sums_tolerance(gpu, ref, rel=1e-5):
    """Per-entry check of reduced [J|r]^T[J|r] sums (28 upper-triangle
    entries + count).  Entry (i,j) is a sum of products J_i J_j whose absolute
    terms add up to at most sqrt(H_ii H_jj) (Cauchy-Schwarz), which bounds any
    FP32 summation-order difference; rel=1e-5 is ~170 ulp of that bound.
    The count must match exactly."""
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    idx = [(i, j) for i in range(7) for j in range(i, 7)]
    diag = np.zeros(7)
    for k, (i, j) in enumerate(idx):
        if i == j:
            diag[i] = abs(ref[k])
    ok = np.zeros(29, bool)
    for k, (i, j) in enumerate(idx):
        scale = np.sqrt(diag[i] * diag[j])
        ok[k] = abs(gpu[k] - ref[k]) <= rel * scale + 1e-12
    ok[28] = gpu[28] == ref[28]
    return ok
