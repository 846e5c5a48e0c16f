"""The Gauss-Newton step's LDLT with its pivots found up front
(ldlt_solve6_presorted in yetanotherslam_amd/csrc/yas_tracker_dev.hpp)
against the oracle's Eigen 3.3 LDLT, which pivots as it factorises
(oracle/oracle.cpp ldlt_solve6).  Eigen's pivot k is the largest |diagonal|
among positions k..5 of the matrix permuted so far, and those entries are
still the input's when step k searches them, so both orders must give the
same bits.  The device code is compiled for the host here (hipcc), on
float-valued systems like the tracker's: random Gram matrices, indefinite
ones, tied and zero diagonals, zero rows, NaN entries and wide magnitudes."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "yetanotherslam_amd", "csrc", "yas_tracker_dev.hpp")
HIPCC = "/opt/rocm/bin/hipcc"

WRAPPER = r'''
#include "%s"
extern "C" void presorted(const float* H, const float* b, double* x) {
  double out[6], scratch[6];
  yas::ldlt_solve6_presorted(H, b, scratch, out);
  for (int i = 0; i < 6; ++i) x[i] = out[i];
}
'''


@pytest.fixture(scope="module")
def presorted(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("ldlt")
    src = d / "w.hip"
    src.write_text(WRAPPER % os.path.abspath(HEADER))
    so = d / "w.so"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950", "-fPIC", "-shared",
                    str(src), "-o", str(so)], check=True, capture_output=True)
    lib = ctypes.CDLL(str(so))
    lib.presorted.argtypes = [ctypes.c_void_p] * 3

    def run(H, b):
        H = np.ascontiguousarray(H, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        x = np.zeros(6, np.float64)
        lib.presorted(H.ctypes.data, b.ctypes.data, x.ctypes.data)
        return x
    yield run
    shutil.rmtree(d, ignore_errors=True)


def systems(n, seed=7):
    rng = np.random.default_rng(seed)
    for c in range(n):
        kind = c % 8
        if kind <= 1:
            J = rng.uniform(-1, 1, (3 + c % 8, 6)).astype(np.float32) * (1e3 if kind else 1)
            H = (J.T @ J).astype(np.float32)
        else:
            M = rng.uniform(-1, 1, (6, 6)).astype(np.float32)
            H = np.tril(M) + np.tril(M, -1).T
        if kind == 3:  # tied diagonal magnitudes
            v = np.float32(rng.uniform(-1, 1))
            for i in range(6):
                if rng.random() < 0.5:
                    H[i, i] = v if rng.random() < 0.5 else -v
        if kind == 4:  # zero diagonal entries and rows
            for i in range(6):
                if rng.random() < 0.35:
                    H[i, i] = 0
                    if rng.random() < 0.5:
                        H[i, :] = 0
                        H[:, i] = 0
        if kind == 5 and c % 40 == 5:
            H[:] = 0
        if kind == 6:  # a NaN
            i, j = rng.integers(0, 6, 2)
            H[i, j] = H[j, i] = np.nan
        if kind == 7:  # wide magnitudes, kept symmetric
            S = np.ldexp(np.float32(1), rng.integers(-20, 20, 6)).astype(np.float32)
            H = (H * S[:, None] * S[None, :]).astype(np.float32)
        b = rng.uniform(-1, 1, 6).astype(np.float32)
        yield H.astype(np.float32), b


def test_presorted_ldlt_is_bit_identical_to_eigen_order(presorted):
    n = 0
    for H, b in systems(20000):
        got = presorted(H, b)
        want = po.ldlt_solve6(H.astype(np.float64), b.astype(np.float64))
        assert np.array_equal(got.view(np.uint64), np.asarray(want, np.float64).view(np.uint64)), (H, b, got, want)
        n += 1
    assert n == 20000
