"""ctypes bindings of the scalar C++ oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker and the CPU baseline,
never the thing measured or shipped.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

VOX_REF, VOX_F32, VOX_F16 = 0, 1, 2
MODE_RGB, MODE_DEPTH, MODE_RGBD = 0, 1, 2

LIB_OMP_PATH = os.path.join(_HERE, "_build", "liboracle_omp.so")

_lib = None
_lib_omp = None
P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH)
    return _lib


def lib_omp():
    """The OpenMP build (all host cores), for timing the CPU baseline only."""
    global _lib_omp
    if _lib_omp is None:
        _lib_omp = _load(LIB_OMP_PATH)
    return _lib_omp


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    sig = {
        "orc_map_new": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                            ctypes.c_double, ctypes.c_double]),
        "orc_map_free": (None, [P]),
        "orc_map_fuse": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, P, P]),
        "orc_map_allocate": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, P, P]),
        "orc_map_visible": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int]),
        "orc_map_num_blocks": (ctypes.c_int, [P]),
        "orc_map_fuse_map": (ctypes.c_int, [P, P, P, P]),
        "orc_map_export": (ctypes.c_int, [P, P, P, P, ctypes.c_int]),
        "orc_map_import": (None, [P, P, P, P, ctypes.c_int]),
        "orc_map_raycast": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, P, P, P, P]),
        "orc_map_mesh": (ctypes.c_long, [P, P, P, P, P, P, ctypes.c_long, ctypes.c_int]),
        "orc_ref_hash": (ctypes.c_int, [ctypes.c_int] * 4),
        "orc_threads": (ctypes.c_int, []),
        "orc_vox_to_block": (ctypes.c_int, [ctypes.c_int]),
        "orc_vox_to_local": (ctypes.c_int, [ctypes.c_int]),
        "orc_pack_sdf": (ctypes.c_int16, [ctypes.c_float]),
        "orc_f32_to_f16": (ctypes.c_uint16, [ctypes.c_float]),
        "orc_f16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
        "orc_mc_table": (ctypes.POINTER(ctypes.c_uint64), []),
        "orc_mc_cube": (ctypes.c_int, [ctypes.c_int, P, P, ctypes.c_double, P, P, ctypes.POINTER(ctypes.c_int)]),
        "orc_se3_exp": (None, [P, P]),
        "orc_ldlt_solve6": (None, [P, P, P]),
        "orc_image_op": (None, [ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P, P]),
        "orc_trk_new": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P]),
        "orc_trk_free": (None, [P]),
        "orc_trk_set_ref_image": (None, [P, P]),
        "orc_trk_set_ref_depth": (None, [P, P]),
        "orc_trk_set_cur_image": (None, [P, P]),
        "orc_trk_set_cur_depth": (None, [P, P]),
        "orc_trk_set_ref_model": (None, [P, P]),
        "orc_trk_step_sums": (None, [P, ctypes.c_int, P, ctypes.c_int, P, P]),
        "orc_trk_get_transform": (ctypes.c_int, [P, P, ctypes.c_int, P]),
        "orc_trk_swap": (None, [P]),
        "orc_trk_get_level": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P]),
        "orc_depth_u16_to_f32": (None, [P, P, ctypes.c_long, ctypes.c_double]),
        "orc_rgb_to_intensity": (None, [P, P, ctypes.c_long, ctypes.c_int]),
        "orc_pipe_new": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                             ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P,
                             ctypes.c_double, ctypes.c_int]),
        "orc_pipe_free": (None, [P]),
        "orc_pipe_map": (P, [P]),
        "orc_pipe_tracker": (P, [P]),
        "orc_pipe_step": (ctypes.c_int, [P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _p(a):
    return a.ctypes.data if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleMap:
    """Voxel-hashed TSDF volume (unordered_map of 8^3 blocks)."""

    def __init__(self, voxel, trunc, precision=32, voxfmt=VOX_F32, max_weight=255, dmin=0.1, dmax=3.0):
        self.h = lib().orc_map_new(precision, voxfmt, voxel, trunc, max_weight, dmin, dmax)
        self.precision = precision

    def __del__(self):
        # at interpreter exit the module's globals may already be gone
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_map_free(self.h)
            self.h = None

    def fuse(self, depth_f32, K, T):
        d = np.ascontiguousarray(depth_f32, np.float32)
        return lib().orc_map_fuse(self.h, _p(d), d.shape[1], d.shape[0], _p(_f64(K)), _p(_f64(T)))

    def allocate(self, depth_f32, K, T):
        d = np.ascontiguousarray(depth_f32, np.float32)
        return lib().orc_map_allocate(self.h, _p(d), d.shape[1], d.shape[0], _p(_f64(K)), _p(_f64(T)))

    def visible(self, W, H, K, T):
        n = self.num_blocks()
        keys = np.zeros((max(n, 1), 3), np.int32)
        m = lib().orc_map_visible(self.h, W, H, _p(_f64(K)), _p(_f64(T)), _p(keys), n)
        return keys[:m]

    def num_blocks(self):
        return lib().orc_map_num_blocks(self.h)

    def fuse_map(self, src, T_self, T_src):
        """Merge `src` (same voxel size and storage) into this map; T_self and
        T_src are the two maps' poses (map -> world).  Returns the number of
        blocks created here."""
        r = lib().orc_map_fuse_map(self.h, src.h, _p(_f64(T_self)), _p(_f64(T_src)))
        if r < 0:
            raise ValueError("maps differ in storage or precision")
        return r

    def export(self, with_voxels=True):
        n = self.num_blocks()
        keys = np.zeros((max(n, 1), 3), np.int32)
        sdf = np.zeros((max(n, 1), 512), np.float64) if with_voxels else None
        wt = np.zeros((max(n, 1), 512), np.float64) if with_voxels else None
        lib().orc_map_export(self.h, _p(keys), _p(sdf), _p(wt), n)
        if with_voxels:
            return keys[:n], sdf[:n], wt[:n]
        return keys[:n]

    def import_blocks(self, keys, sdf, wt):
        """Blocks (n, 3) int32 keys with (n, 512) decoded sdf and weight
        values, as export gives them (test set-up for the read-only
        functions on a volume built elsewhere)."""
        k = np.ascontiguousarray(keys, np.int32)
        s = np.ascontiguousarray(sdf, np.float64)
        w = np.ascontiguousarray(wt, np.float64)
        assert k.shape[0] == s.shape[0] == w.shape[0] and s.shape[1] == w.shape[1] == 512
        lib().orc_map_import(self.h, _p(k), _p(s), _p(w), k.shape[0])

    def raycast(self, W, H, K, T):
        vmap = np.zeros((H, W, 4), np.float64)
        zmin = np.zeros((H // 8, W // 8), np.float64)
        zmax = np.zeros((H // 8, W // 8), np.float64)
        if lib().orc_map_raycast(self.h, W, H, _p(_f64(K)), _p(_f64(T)), _p(vmap), _p(zmin), _p(zmax)) < 0:
            raise ValueError("raycast: width and height must be multiples of 8, got %dx%d" % (W, H))
        return vmap, zmin, zmax

    def mesh(self, tri_cap=None):
        nb = self.num_blocks()
        if tri_cap is None:
            tri_cap = lib().orc_map_mesh(self.h, None, None, None, None, None, 0, 0)
        keys = np.zeros((max(nb, 1), 3), np.int32)
        counts = np.zeros(max(nb, 1), np.int32)
        cases = np.zeros((max(nb, 1), 512), np.int16)
        tris = np.zeros((max(tri_cap, 1), 9), np.float64)
        normals = np.zeros((max(tri_cap, 1), 9), np.float64)
        n = lib().orc_map_mesh(self.h, _p(keys), _p(counts), _p(cases), _p(tris), _p(normals), tri_cap, nb)
        return dict(keys=keys[:nb], counts=counts[:nb], cases=cases[:nb], tris=tris[:n], normals=normals[:n])


def image_op(op, src, precision=32, params=None):
    s = np.ascontiguousarray(src, np.float32)
    H, W = s.shape[0], s.shape[1]
    if op == 0:
        out = np.zeros((H // 2, W // 2), np.float64)
    elif op == 1:
        out = np.zeros((2, H, W), np.float64)
    elif op == 2:
        out = np.zeros((H, W), np.float64)
    else:
        out = np.zeros((H, W, 4), np.float64)
    p = _f64(params if params is not None else [0, 0, 0, 0])
    lib().orc_image_op(op, precision, _p(s), W, H, _p(out), _p(p))
    return out


def pyrdown_gauss(img, precision=32):
    return image_op(0, img, precision)


def gradient(img, precision=32):
    g = image_op(1, img, precision)
    return g[0], g[1]


def depth_to_invdepth(d, precision=32):
    return image_op(2, d, precision)


def vertex_map(inv, invfx, invfy, cx, cy, precision=32):
    return image_op(3, inv, precision, [invfx, invfy, cx, cy])


def normal_map(vmap, precision=32):
    return image_op(4, vmap, precision)


def se3_exp(xi):
    T = np.zeros((4, 4), np.float64)
    lib().orc_se3_exp(_p(_f64(xi)), _p(T))
    return T


def ldlt_solve6(A, b):
    x = np.zeros(6, np.float64)
    lib().orc_ldlt_solve6(_p(_f64(A)), _p(_f64(b)), _p(x))
    return x


def mc_table():
    t = lib().orc_mc_table()
    return np.array([t[i] for i in range(256)], dtype=np.uint64)


def mc_cube(p, s, vs=1.0, precision=64):
    """One marching cube: (cube index, triangles (n, 9), normals (n, 9))."""
    pp = np.ascontiguousarray(p, np.int32)
    ss = _f64(s)
    tris = np.zeros(45, np.float64)
    nrm = np.zeros(45, np.float64)
    n = ctypes.c_int()
    c = lib().orc_mc_cube(precision, _p(pp), _p(ss), vs, _p(tris), _p(nrm), ctypes.byref(n))
    return c, tris[:9 * n.value].reshape(-1, 9), nrm[:9 * n.value].reshape(-1, 9)


def depth_u16_to_f32(d, scale=1000.0):
    d = np.ascontiguousarray(d, np.uint16)
    z = np.zeros(d.shape, np.float32)
    lib().orc_depth_u16_to_f32(_p(d), _p(z), d.size, scale)
    return z


def rgb_to_intensity(rgb, mode=0):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    g = np.zeros(rgb.shape[:2], np.float32)
    lib().orc_rgb_to_intensity(_p(rgb), _p(g), g.size, mode)
    return g


class OracleTracker:
    def __init__(self, W, H, K, levels=3, iters=(10, 5, 3), mode=MODE_DEPTH, weights=(1, 1, 100, 10),
                 precision=32):
        self.W, self.H, self.levels = W, H, levels
        it = np.array(list(iters) + [0] * (8 - len(iters)), np.int32)
        w = np.array(weights, np.float32)
        self.h = lib().orc_trk_new(precision, W, H, _p(_f64(K)), levels, _p(it), mode, _p(w))

    def __del__(self):
        # at interpreter exit the module's globals may already be gone
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_trk_free(self.h)
            self.h = None

    def _f32(self, a):
        return np.ascontiguousarray(a, np.float32)

    def set_reference_image(self, g):
        a = self._f32(g); lib().orc_trk_set_ref_image(self.h, _p(a))

    def set_reference_depth(self, d):
        a = self._f32(d); lib().orc_trk_set_ref_depth(self.h, _p(a))

    def set_tracking_image(self, g):
        a = self._f32(g); lib().orc_trk_set_cur_image(self.h, _p(a))

    def set_tracking_depth(self, d):
        a = self._f32(d); lib().orc_trk_set_cur_depth(self.h, _p(a))

    def set_reference_model(self, vmap):
        a = self._f32(vmap); lib().orc_trk_set_ref_model(self.h, _p(a))

    def step_sums(self, lvl, T, kind):
        out = np.zeros(29, np.float64)
        stats = np.zeros(4, np.float64)
        lib().orc_trk_step_sums(self.h, lvl, _p(_f64(T)), kind, _p(out), _p(stats))
        return out, stats

    def get_transform(self, init=None, swap=True):
        init = np.eye(4) if init is None else init
        out = np.zeros((4, 4), np.float64)
        good = lib().orc_trk_get_transform(self.h, _p(_f64(init)), 1 if swap else 0, _p(out))
        return out, bool(good)

    def swap(self):
        lib().orc_trk_swap(self.h)

    def level(self, which, lvl):
        w, h = self.W >> lvl, self.H >> lvl
        ch = 1 if which < 6 else 4
        out = np.zeros((h, w, ch) if ch == 4 else (h, w), np.float64)
        lib().orc_trk_get_level(self.h, which, lvl, _p(out))
        return out


class OraclePipeline:
    """Tracking::TrackRGBD replayed on one volume (see oracle.cpp).
    all_cores: run on the OpenMP build (timing only; its sums add in another
    order)."""

    def __init__(self, W, H, K, voxel, trunc, levels, iters, mode, weights=(1, 1, 100, 10), precision=32,
                 voxfmt=VOX_F32, max_weight=255, depth_scale=1000.0, gray_mode=0, all_cores=False):
        it = np.array(list(iters) + [0] * (8 - len(iters)), np.int32)
        w = np.array(weights, np.float32)
        self.L = lib_omp() if all_cores else lib()
        self.h = self.L.orc_pipe_new(precision, voxfmt, voxel, trunc, max_weight, W, H, _p(_f64(K)), levels,
                                     _p(it), mode, _p(w), depth_scale, gray_mode)
        if not self.h:
            raise ValueError("pipeline: width and height must be multiples of 8, got %dx%d" % (W, H))
        self.W, self.H = W, H

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "L", None) is not None:
            self.L.orc_pipe_free(self.h)
            self.h = None

    def step(self, depth_u16, rgb):
        d = np.ascontiguousarray(depth_u16, np.uint16)
        c = np.ascontiguousarray(rgb, np.uint8)
        T = np.zeros((4, 4), np.float64)
        times = np.zeros(4, np.float64)
        ok = self.L.orc_pipe_step(self.h, _p(d), _p(c), _p(T), _p(times))
        return T, bool(ok), times
