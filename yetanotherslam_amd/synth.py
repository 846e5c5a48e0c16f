"""Synthetic RGB-D sequences (room + spheres, circular camera path).

Thin ctypes wrapper over ``libyas_synth.so`` (built from ``csrc/synth.cpp``).
It replaces the reference's dataset/camera readers
(Examples/tum_rgbd.cpp:43-57, Thirdparty/OpenNI2/src/Camera.cpp:42-51) as the
source of frames for tests and benchmarks.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class SynthParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("fx", ctypes.c_double), ("fy", ctypes.c_double),
        ("cx", ctypes.c_double), ("cy", ctypes.c_double),
        ("seed", ctypes.c_uint32), ("noise_seed", ctypes.c_uint32),
        ("theta0", ctypes.c_double), ("step", ctypes.c_double),
        ("radius", ctypes.c_double), ("cam_height", ctypes.c_double),
        ("noise_k", ctypes.c_double), ("dropout", ctypes.c_double),
        ("depth_scale", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libyas_synth.so")
        if not os.path.exists(path):
            raise RuntimeError("libyas_synth.so is not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(path)
        _lib.yas_synth_frame.restype = ctypes.c_int
        _lib.yas_synth_frame.argtypes = [ctypes.POINTER(SynthParams), ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.yas_synth_pose.argtypes = [ctypes.POINTER(SynthParams), ctypes.c_int, ctypes.c_void_p]
        _lib.yas_synth_spheres.argtypes = [ctypes.POINTER(SynthParams), ctypes.c_void_p]
        _lib.yas_synth_render.restype = ctypes.c_int
        _lib.yas_synth_render.argtypes = [ctypes.POINTER(SynthParams), ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return _lib


def params(width=640, height=480, seed=0, noise_seed=0, theta0=0.0, step=0.02, radius=1.0,
           cam_height=1.5, noise_k=0.0012, dropout=0.02, depth_scale=1000.0):
    """Default camera: fx=fy=525 at 640x480, scaled for other resolutions
    (principal point (c+0.5)*s-0.5)."""
    p = SynthParams()
    s = width / 640.0
    p.width, p.height = width, height
    p.fx = p.fy = 525.0 * s
    p.cx = (319.5 + 0.5) * s - 0.5
    p.cy = (239.5 + 0.5) * (height / 480.0) - 0.5
    p.seed, p.noise_seed = seed, noise_seed
    p.theta0, p.step, p.radius, p.cam_height = theta0, step, radius, cam_height
    p.noise_k, p.dropout, p.depth_scale = noise_k, dropout, depth_scale
    return p


def intrinsics(p):
    return np.array([p.fx, p.fy, p.cx, p.cy], dtype=np.float64)


def frame(p, t, clean=False):
    """Returns (depth_u16 HxW, rgb_u8 HxWx3, T_world_cam 4x4[, clean depth f32])."""
    H, W = p.height, p.width
    depth = np.zeros((H, W), np.uint16)
    rgb = np.zeros((H, W, 3), np.uint8)
    cd = np.zeros((H, W), np.float32) if clean else None
    T = np.zeros((4, 4), np.float64)
    rc = lib().yas_synth_frame(ctypes.byref(p), int(t), depth.ctypes.data, rgb.ctypes.data,
                               cd.ctypes.data if clean else None, T.ctypes.data)
    if rc != 0:
        raise RuntimeError("yas_synth_frame failed")
    if clean:
        return depth, rgb, T, cd
    return depth, rgb, T


def frame_at(p, T_world_cam, t):
    """The scene of `p` from camera pose T_world_cam (4x4); `t` selects the
    noise draws.  Returns (depth_u16 HxW, rgb_u8 HxWx3)."""
    H, W = p.height, p.width
    depth = np.zeros((H, W), np.uint16)
    rgb = np.zeros((H, W, 3), np.uint8)
    T = np.ascontiguousarray(T_world_cam, np.float64)
    if lib().yas_synth_render(ctypes.byref(p), T.ctypes.data, int(t), depth.ctypes.data, rgb.ctypes.data, None):
        raise RuntimeError("yas_synth_render failed")
    return depth, rgb


def wobble_pose(p, t, pitch=0.05, lift=0.03, roll=0.0):
    """The circular path's pose at frame t with the camera pitched by
    pitch*sin(t/2) rad about its x axis, rolled by roll*t rad about its
    optical axis and raised by lift*sin(t/3) m: motions between frames no
    longer share one screw axis, so they do not commute."""
    T = pose(p, t)
    a = pitch * np.sin(t / 2.0)
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    b = roll * t
    Rz = np.array([[np.cos(b), -np.sin(b), 0], [np.sin(b), np.cos(b), 0], [0, 0, 1]])
    out = T.copy()
    out[:3, :3] = T[:3, :3] @ Rx @ Rz
    out[1, 3] += lift * np.sin(t / 3.0)
    return out


def pose(p, t):
    T = np.zeros((4, 4), np.float64)
    lib().yas_synth_pose(ctypes.byref(p), int(t), T.ctypes.data)
    return T


def spheres(p):
    out = np.zeros((3, 4), np.float64)
    lib().yas_synth_spheres(ctypes.byref(p), out.ctypes.data)
    return out


def relative_poses(p, n):
    """Ground-truth camera-to-map poses with the map frame = camera 0."""
    T0inv = np.linalg.inv(pose(p, 0))
    return [T0inv @ pose(p, t) for t in range(n)]
