"""The Gauss-Newton step's 6x6 solve as the tracker's kernels run it
(ldlt_solve6_wave in yetanotherslam_amd/csrc/yas_tracker_dev.hpp: one
wavefront, one matrix row per lane) against the oracle's Eigen 3.3 LDLT
(oracle/oracle.cpp ldlt_solve6, RGBDTracking.cu:185), bit for bit, on the
systems of test_ldlt_presorted: Gram matrices, indefinite ones, tied and
zero diagonals, zero rows, NaN entries and wide magnitudes.  A NaN result
must be NaN on both sides (the GPU's and the host's NaN payloads differ)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yetanotherslam_amd import rgbdfusion as rf
from yetanotherslam_amd.native import YasError, call

from test_ldlt_presorted import systems


@pytest.mark.gpu
def test_wave_solve_is_bit_identical_to_eigen_order(gpu):
    S = list(systems(20000))
    H = np.ascontiguousarray(np.stack([h for h, _ in S]), np.float32)
    b = np.ascontiguousarray(np.stack([v for _, v in S]), np.float32)
    x = np.full((len(S), 6), -1.0, np.float64)
    call("yas_tracker_solve6", rf.default_context(0).h, H.ctypes.data, b.ctypes.data, x.ctypes.data, len(S))
    want = np.stack([np.asarray(po.ldlt_solve6(h.astype(np.float64), v.astype(np.float64)), np.float64)
                     for h, v in S])
    same = (x.view(np.uint64) == want.view(np.uint64)) | (np.isnan(x) & np.isnan(want))
    bad = np.flatnonzero(~same.all(1))
    print("systems", len(S), "with a NaN solution", int(np.isnan(want).any(1).sum()), "differing", bad.size)
    assert bad.size == 0, (bad[:5], H[bad[0]], b[bad[0]], x[bad[0]], want[bad[0]])


@pytest.mark.gpu
def test_wave_solve_refuses_bad_counts(gpu):
    z = np.zeros(36, np.float32)
    x = np.zeros(6, np.float64)
    ctx = rf.default_context(0).h
    for n in (0, -1, (1 << 20) + 1):
        with pytest.raises(YasError):
            call("yas_tracker_solve6", ctx, z.ctypes.data, z.ctypes.data, x.ctypes.data, n)
