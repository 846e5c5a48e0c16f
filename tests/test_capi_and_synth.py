"""CPU checks of the boundary and the input generator: libyas.so loads and
exports every entry point include/yas.h declares (no compute call is made
without a GPU), the Python bindings cover them, the product has no CPU
fallback, and the synthetic sequence is deterministic and geometrically
right."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from yetanotherslam_amd import native, synth

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    with open(os.path.join(REPO, "include", "yas.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void\*)\s+(yas_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    lib = ctypes.CDLL(native.LIB_PATH)
    names = declared_entry_points()
    assert len(names) > 50
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(native.SIGNATURES), set(names) ^ set(native.SIGNATURES)


def test_version_and_error_calls_need_no_device():
    lib = native.lib()
    assert b"gfx950" in lib.yas_version()
    # an invalid argument is rejected on the host with a message, no device touched
    rc = lib.yas_map_create(None, None, None, None)
    assert rc == 1
    assert b"null" in lib.yas_last_error()


def test_product_has_no_cpu_fallback(monkeypatch, tmp_path):
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(native, "_lib", None)
    with pytest.raises(RuntimeError):
        native.lib()


def test_synthetic_frames_are_deterministic():
    p = synth.params(width=160, height=120)
    a = synth.frame(p, 7)
    b = synth.frame(p, 7)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    c = synth.frame(synth.params(width=160, height=120, noise_seed=1), 7)
    assert not np.array_equal(a[0], c[0])


def test_camera_path_and_depth_geometry():
    p = synth.params(width=160, height=120, theta0=math.pi / 2, noise_k=0.0, dropout=0.0)
    for t in (0, 10, 50):
        T = synth.pose(p, t)
        R = T[:3, :3]
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(R) - 1) < 1e-12
        c = T[:3, 3]
        assert abs(math.hypot(c[0], c[2]) - 1.0) < 1e-12 and abs(c[1] - 1.5) < 1e-12
        fwd = R[:, 2]
        to_axis = np.array([-c[0], 0.0, -c[2]]) / math.hypot(c[0], c[2])
        np.testing.assert_allclose(fwd, to_axis, atol=1e-12)  # looking at the room's axis
    # theta0 = pi/2: camera at (0, 1.5, 1) looking down -z at the wall z = -1.5
    d, rgb, T, clean = synth.frame(p, 0, clean=True)
    sph = synth.spheres(p)
    cy, cx = p.height // 2, p.width // 2
    ray = T[:3, :3] @ np.array([(cx - p.cx) / p.fx, (cy - p.cy) / p.fy, 1.0])
    o = T[:3, 3]
    hits = [(-1.5 - o[2]) / ray[2]]
    for s in sph:
        oc = o - s[:3]
        B = 2 * oc @ ray
        C = oc @ oc - s[3] ** 2
        disc = B * B - 4 * (ray @ ray) * C
        if disc >= 0:
            tt = (-B - math.sqrt(disc)) / (2 * (ray @ ray))
            if tt > 0:
                hits.append(tt)
    assert abs(clean[cy, cx] - min(hits)) < 1e-5
    assert abs(d[cy, cx] - round(1000 * min(hits))) <= 1


def test_depth_noise_and_dropout_statistics():
    p = synth.params(width=320, height=240, theta0=math.pi / 4)
    d, _, _, clean = synth.frame(p, 3, clean=True)
    drop = d == 0
    assert 0.015 < drop.mean() < 0.03
    ok = ~drop & (clean < 3.0)
    z = d[ok] / 1000.0
    resid = (z - clean[ok]) / (0.0012 * clean[ok] ** 2)
    assert abs(resid.mean()) < 0.1 and 0.8 < resid.std() < 1.3  # N(0, (k z^2)^2) plus mm rounding


def test_cpp_facade_compiles_and_links(tmp_path):
    """The reference-shaped C++ facade (include/yas_rgbdfusion.hpp) compiles
    against yas.h and links against libyas.so; the GPU test runs it."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = tmp_path / "facade_replay"
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O1", "-std=c++17", os.path.join(REPO, "tests", "cpp",
                        "facade_replay.cpp"), "-L" + os.path.dirname(native.LIB_PATH), "-lyas",
                        "-Wl,-rpath," + os.path.dirname(native.LIB_PATH), "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
