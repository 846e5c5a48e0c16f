"""Python mirror of the reference's dense-fusion classes over the C ABI.

Class and method names follow RGBDFusion/include (MapStruct, RayTraceEngine,
MeshEngine, RGBDTracking and the ImageProc free functions) so that tests read
like calls into the reference.  Device images are torch tensors on the
context's device (the reference passes cv::cuda::GpuMat); poses are 4x4 numpy
arrays (the reference passes Sophus::SE3d).  A Context runs the library's work
on the torch stream that is current when it is made: a stream torch created
is shared, and torch's default stream (handle 0) is the device's null stream,
which the library then binds to explicitly.  Torch work and library work are
therefore ordered on one stream, and torch's caching allocator, which reuses
a freed block only in that stream's order, cannot hand an input still being
read to a later copy.

Every native handle is closed in dependency order (servers and batches,
pipelines, engines, maps, contexts) by an at-exit hook that runs before the
interpreter starts tearing modules down, so no HIP call is made during
interpreter finalisation.

PyTorch is plumbing here (device memory and streams); every computation runs
in libyas.so's HIP kernels, and nothing falls back to the CPU.
"""
import atexit
import ctypes
import weakref

import numpy as np
import torch

from . import native
from .native import BatchConfig, MapParams, MapStats, Result, ServerStats, TrackerParams, TrackStats, call

VOXEL_I16_U8, VOXEL_F32_F32, VOXEL_F16_U16 = 0, 1, 2
TRACK_RGB, TRACK_DEPTH, TRACK_RGBD = 0, 1, 2
HIP_MEMCPY_D2H = 2

_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        native.lib()
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


def download(ptr, shape, dtype):
    """Copy device memory at `ptr` into a new numpy array (synchronous)."""
    out = np.empty(shape, dtype)
    rc = _hiprt().hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, HIP_MEMCPY_D2H)
    if rc != 0:
        raise RuntimeError("hipMemcpy failed: %d" % rc)
    return out


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _dev(t):
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError("expected a CUDA/HIP torch tensor")
    return ctypes.c_void_p(t.data_ptr())


def _pitch(t):
    return t.stride(0) * t.element_size()


# Live native handles by teardown rank: 0 servers and batches, 1 pipelines,
# 2 trackers, ray casters and mesh engines, 3 maps, 4 contexts.
_live = [weakref.WeakSet() for _ in range(5)]
_exiting = False


def close_all():
    """Close every live handle, dependents before what they depend on."""
    for rank in _live:
        for obj in list(rank):
            try:
                obj.close()
            except Exception:
                pass


@atexit.register
def _close_at_exit():
    global _exiting
    close_all()
    _exiting = True


class _Native:
    """A library handle `self.h`, destroyed by `_DESTROY` on close()."""
    _DESTROY = None
    _RANK = 3

    def _own(self, h):
        self.h = h
        _live[self._RANK].add(self)

    def close(self):
        if getattr(self, "h", None):
            getattr(native.lib(), self._DESTROY)(self.h)
            self.h = None
        _live[self._RANK].discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if _exiting:
            return
        try:
            self.close()
        except Exception:
            pass


class Context(_Native):
    """yas_ctx on `device`, bound to torch's current stream of that device
    (the null stream when that is torch's default stream)."""
    _DESTROY = "yas_ctx_destroy"
    _RANK = 4

    def __init__(self, device=0):
        self.device = device
        with torch.cuda.device(device):
            stream = torch.cuda.current_stream(device).cuda_stream
        h = ctypes.c_void_p()
        if stream:
            call("yas_ctx_create", device, ctypes.c_void_p(stream), ctypes.byref(h))
        else:
            call("yas_ctx_create_null_stream", device, ctypes.byref(h))
        self._own(h)
        self.stream_handle = stream

    def synchronize(self):
        call("yas_ctx_synchronize", self.h)


_default_ctx = {}


def default_context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def decode_voxels(raw, fmt):
    """Raw exported voxels (n, 512) -> (sdf float32, weight float32)."""
    if fmt == VOXEL_F32_F32:
        a = raw.view(np.float32).reshape(raw.shape[0], 512, 2)
        return a[..., 0].copy(), a[..., 1].copy()
    v = raw.view(np.uint32).reshape(raw.shape[0], 512)
    lo = (v & 0xFFFF).astype(np.uint16)
    if fmt == VOXEL_I16_U8:
        sdf = lo.view(np.int16).astype(np.float32) / np.float32(32767.0)
        wt = ((v >> 16) & 0xFF).astype(np.float32)
    else:
        sdf = lo.view(np.float16).astype(np.float32)
        wt = (v >> 16).astype(np.float32)
    return sdf, wt


class MapStruct(_Native):
    """MapStruct (RGBDFusion/include/VoxelMap.h:32-144)."""
    _DESTROY = "yas_map_destroy"
    _RANK = 3

    def __init__(self, K, num_buckets, num_blocks, voxel_size, trunc_dist, voxel_format=VOXEL_F32_F32,
                 max_weight=255, depth_min=0.1, depth_max=3.0, ctx=None):
        self.ctx = ctx or default_context()
        p = MapParams(num_buckets, num_blocks, voxel_size, trunc_dist, voxel_format, max_weight, depth_min, depth_max)
        self.params = p
        self.K = np.asarray(K, np.float32)
        h = ctypes.c_void_p()
        call("yas_map_create", self.ctx.h, ctypes.byref(p), _ptr(self.K), ctypes.byref(h))
        self._own(h)
        self.voxel_format = voxel_format

    def Reset(self):
        call("yas_map_reset", self.h)

    def SetPose(self, T):
        call("yas_map_set_pose", self.h, _ptr(_f64(T)))

    def GetPose(self):
        T = np.zeros((4, 4))
        call("yas_map_get_pose", self.h, _ptr(T))
        return T

    def Fuse(self, depth, T):
        """Fuse(GpuMat depth_f32_metres, SE3d Tcm) (VoxelMap.cu:825-884)."""
        assert depth.dtype == torch.float32 and depth.dim() == 2
        call("yas_map_fuse", self.h, _dev(depth), _pitch(depth), depth.shape[1], depth.shape[0], _ptr(_f64(T)))

    def FuseU16(self, depth, T, depth_scale=1000.0):
        assert depth.dtype in (torch.int16, torch.uint16) and depth.dim() == 2
        call("yas_map_fuse_u16", self.h, _dev(depth), _pitch(depth), depth.shape[1], depth.shape[0],
             depth_scale, _ptr(_f64(T)))

    def CheckNumVisibleBlocks(self, cols, rows, T):
        n = ctypes.c_int()
        call("yas_map_check_visible_blocks", self.h, cols, rows, _ptr(_f64(T)), ctypes.byref(n))
        return n.value

    def visible_blocks(self):
        """Keys (n, 3) of the blocks the last visibility pass found visible."""
        n = self.stats().num_visible
        keys = np.zeros((max(n, 1), 3), np.int32)
        cnt = ctypes.c_int()
        call("yas_map_visible_blocks", self.h, _ptr(keys), n, ctypes.byref(cnt))
        return keys[:cnt.value]

    def stats(self):
        s = MapStats()
        call("yas_map_get_stats", self.h, ctypes.byref(s))
        return s

    def export(self):
        """(keys int32 (n,3), raw voxels (n, 512*bytes) uint8) in pool order."""
        n = self.stats().num_allocated
        vb = native.lib().yas_map_voxel_bytes(self.h)
        keys = np.zeros((max(n, 1), 3), np.int32)
        raw = np.zeros((max(n, 1), 512 * vb), np.uint8)
        cnt = ctypes.c_int()
        call("yas_map_export", self.h, _ptr(keys), _ptr(raw), n, ctypes.byref(cnt))
        return keys[:n], raw[:n]

    def export_sorted(self):
        """Blocks sorted by (x, y, z): keys, sdf (n,512) float32, weight (n,512) float32."""
        keys, raw = self.export()
        order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
        sdf, wt = decode_voxels(raw[order], self.voxel_format)
        return keys[order], sdf, wt

    def FuseMap(self, src):
        """Fuse(MapStruct*) (VoxelMap.cu:400-449) without releasing src."""
        call("yas_map_fuse_map", self.h, src.h)

    def Reserve(self, num_buckets, num_blocks):
        call("yas_map_reserve", self.h, num_buckets, num_blocks)

    def Hibernate(self):
        call("yas_map_hibernate", self.h)

    def ReActivate(self):
        call("yas_map_reactivate", self.h)

    def is_hibernating(self):
        v = ctypes.c_int()
        call("yas_map_is_hibernating", self.h, ctypes.byref(v))
        return bool(v.value)

    def SaveToFile(self, path):
        call("yas_map_save", self.h, str(path).encode())

    @classmethod
    def ReadFromFile(cls, path, ctx=None):
        ctx = ctx or default_context()
        h = ctypes.c_void_p()
        call("yas_map_load", ctx.h, str(path).encode(), ctypes.byref(h))
        m = cls.__new__(cls)
        m.ctx = ctx
        m._own(h)
        m.params = MapParams()
        m.K = np.zeros(4, np.float32)
        call("yas_map_get_params", h, ctypes.byref(m.params), _ptr(m.K))
        m.voxel_format = m.params.voxel_format
        return m


class _MapView(MapStruct):
    """A MapStruct that borrows the volume of a pipeline or server stream;
    closing it never destroys the volume."""

    def __init__(self, h, voxel_format, owner):
        self.h = h
        self.voxel_format = voxel_format
        self._owner = owner

    def close(self):
        self.h = None


class Server(_Native):
    """Many RGB-D streams tracked and fused with continuous batching
    (yas_server_*)."""
    _DESTROY = "yas_server_destroy"
    _RANK = 0

    def __init__(self, max_width, max_height, max_streams, tracker_params, voxel_format=VOXEL_F32_F32,
                 hbm_budget=200 << 30, ctx=None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        call("yas_server_create", self.ctx.h, max_width, max_height, max_streams, ctypes.byref(tracker_params),
             voxel_format, hbm_budget, ctypes.byref(h))
        self._own(h)
        self._inputs = {}

    def add_stream(self, W, H, K, map_params, depth_scale=1000.0, gray_mode=0):
        Kf = np.asarray(K, np.float32)
        sid = ctypes.c_int()
        call("yas_server_add_stream", self.h, W, H, _ptr(Kf), ctypes.byref(map_params), depth_scale, gray_mode,
             ctypes.byref(sid))
        self._dims = getattr(self, "_dims", {})
        self._dims[sid.value] = (H, W)
        return sid.value

    def open_stream(self, W, H, K, map_params, depth_scale=1000.0, gray_mode=0):
        """A stream admitted as soon as it fits the HBM budget (pending until
        then); frames may be submitted at once."""
        Kf = np.asarray(K, np.float32)
        sid = ctypes.c_int()
        call("yas_server_open_stream", self.h, W, H, _ptr(Kf), ctypes.byref(map_params), depth_scale, gray_mode,
             ctypes.byref(sid))
        self._dims = getattr(self, "_dims", {})
        self._dims[sid.value] = (H, W)
        return sid.value

    def close_stream(self, sid):
        call("yas_server_close_stream", self.h, sid)

    def set_eviction(self, idle_ms):
        call("yas_server_set_eviction", self.h, float(idle_ms))

    STATES = ("free", "pending", "active", "hibernated")

    def stream_state(self, sid):
        v = ctypes.c_int()
        call("yas_server_stream_state", self.h, sid, ctypes.byref(v))
        return self.STATES[v.value]

    def set_gn_group(self, n):
        call("yas_server_set_gn_group", self.h, n)

    def set_max_in_flight(self, n, min_ahead=1):
        call("yas_server_set_max_in_flight", self.h, n, min_ahead)

    def set_max_batch(self, n):
        call("yas_server_set_max_batch", self.h, n)

    def prepare(self):
        """Capture every batch size's graph now instead of at its first batch."""
        call("yas_server_prepare", self.h)

    def remove_stream(self, sid):
        call("yas_server_remove_stream", self.h, sid)

    def submit(self, sid, depth_u16, rgb_u8):
        """Queue a frame (device tensors, ready now, kept until its result is polled)."""
        H, W = self._dims[sid]
        if tuple(depth_u16.shape) != (H, W) or tuple(rgb_u8.shape) != (H, W, 3):
            raise ValueError("frame size differs from the stream's %dx%d" % (W, H))
        req = ctypes.c_longlong()
        call("yas_server_submit", self.h, sid, _dev(depth_u16), _pitch(depth_u16), _dev(rgb_u8), _pitch(rgb_u8),
             ctypes.byref(req))
        self._inputs[req.value] = (depth_u16, rgb_u8)
        return req.value

    def pump(self):
        n = ctypes.c_int()
        call("yas_server_pump", self.h, ctypes.byref(n))
        return n.value

    def poll(self, capacity=256):
        buf = (Result * capacity)()
        n = ctypes.c_int()
        call("yas_server_poll", self.h, buf, capacity, ctypes.byref(n))
        out = []
        for r in buf[:n.value]:
            self._inputs.pop(r.request, None)
            out.append({"stream": r.stream, "request": r.request, "tracked": bool(r.tracked),
                        "pose": np.array(r.pose[:], np.float64).reshape(4, 4)})
        return out

    def drain(self):
        call("yas_server_drain", self.h)

    def stats(self):
        st = ServerStats()
        call("yas_server_get_stats", self.h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def stream_map(self, sid, voxel_format=VOXEL_F32_F32):
        """A non-owning MapStruct view of the stream's volume."""
        m = ctypes.c_void_p()
        call("yas_server_stream_map", self.h, sid, ctypes.byref(m))
        return _MapView(m, voxel_format, self)


class FusionBatch(_Native):
    """Many maps fused together, one launch per stage (yas_batch_*)."""
    _DESTROY = "yas_batch_destroy"
    _RANK = 0

    def __init__(self, maps, ctx=None):
        self.ctx = ctx or default_context()
        self.maps = list(maps)
        arr = (ctypes.c_void_p * len(self.maps))(*[m.h.value for m in self.maps])
        h = ctypes.c_void_p()
        call("yas_batch_create", self.ctx.h, len(self.maps), arr, ctypes.byref(h))
        self._own(h)

    def plan(self, depth, poses, depth_scale=1000.0):
        """depth[s][i]: uint16/int16 device tensor (H, W) or None (map i idles
        at step s); poses[s][i]: 4x4 camera -> map."""
        S, n = len(depth), len(self.maps)
        ptrs = (ctypes.c_void_p * (S * n))()
        pitch = (ctypes.c_size_t * (S * n))()
        W = (ctypes.c_int * (S * n))()
        H = (ctypes.c_int * (S * n))()
        T = np.zeros((S * n, 16))
        for s in range(S):
            for i in range(n):
                k = s * n + i
                d = depth[s][i]
                if d is None:
                    continue
                assert d.dtype in (torch.int16, torch.uint16) and d.dim() == 2 and d.is_contiguous()
                ptrs[k], pitch[k], W[k], H[k] = d.data_ptr(), d.stride(0) * 2, d.shape[1], d.shape[0]
                T[k] = np.asarray(poses[s][i], np.float64).reshape(16)
        self._keep = depth
        call("yas_batch_plan", self.h, S, ptrs, pitch, W, H, depth_scale, _ptr(T))
        self.steps = S

    def autotune(self, reset_maps=True):
        c = BatchConfig()
        call("yas_batch_autotune", self.h, 1 if reset_maps else 0, ctypes.byref(c))
        return c.as_dict()

    def config(self):
        c = BatchConfig()
        call("yas_batch_get_config", self.h, ctypes.byref(c))
        return c.as_dict()

    def set_config(self, **kw):
        c = BatchConfig()
        call("yas_batch_get_config", self.h, ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        call("yas_batch_set_config", self.h, ctypes.byref(c))

    def run(self, first=0, count=None):
        call("yas_batch_run", self.h, first, self.steps - first if count is None else count)


class RayTraceEngine(_Native):
    """RayTraceEngine (RGBDFusion/include/RayTraceEngine.h:11-48)."""
    _DESTROY = "yas_raycaster_destroy"
    _RANK = 2

    def __init__(self, w, h, K, ctx=None):
        self.ctx = ctx or default_context()
        self.w, self.h_ = w, h
        Kf = np.asarray(K, np.float32)
        h_ = ctypes.c_void_p()
        call("yas_raycaster_create", self.ctx.h, w, h, _ptr(Kf), ctypes.byref(h_))
        self._own(h_)

    def RayTrace(self, map_struct, T):
        call("yas_raycast", self.h, map_struct.h, _ptr(_f64(T)))

    def vmap_ptr(self):
        p = ctypes.c_void_p()
        pitch = ctypes.c_size_t()
        call("yas_raycaster_vmap", self.h, ctypes.byref(p), ctypes.byref(pitch))
        return p.value, pitch.value

    def GetVMap(self):
        """Host copy of the last vertex map, (h, w, 4) float32."""
        self.ctx.synchronize()
        p, _ = self.vmap_ptr()
        return download(p, (self.h_, self.w, 4), np.float32)

    def GetZRange(self):
        self.ctx.synchronize()
        zmin, zmax = ctypes.c_void_p(), ctypes.c_void_p()
        cols, rows = ctypes.c_int(), ctypes.c_int()
        call("yas_raycaster_zrange", self.h, ctypes.byref(zmin), ctypes.byref(zmax), ctypes.byref(cols),
             ctypes.byref(rows))
        shape = (rows.value, cols.value)
        return download(zmin.value, shape, np.float32), download(zmax.value, shape, np.float32)

    def GetNumVisibleBlock(self):
        n = ctypes.c_int()
        call("yas_raycaster_num_visible", self.h, ctypes.byref(n))
        return n.value


class MeshEngine(_Native):
    """MeshEngine (RGBDFusion/include/MeshEngine.h:8-24)."""
    _DESTROY = "yas_mesher_destroy"
    _RANK = 2

    def __init__(self, max_triangles, ctx=None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        call("yas_mesher_create", self.ctx.h, max_triangles, ctypes.byref(h))
        self._own(h)
        self.max_triangles = max_triangles

    def Meshify(self, map_struct, with_cases=False):
        n = ctypes.c_int()
        call("yas_mesh_extract", self.h, map_struct.h, ctypes.byref(n), allow=(6,))
        ntri = n.value
        keep = min(ntri, self.max_triangles)
        tris = np.zeros((max(keep, 1), 9), np.float32)
        normals = np.zeros((max(keep, 1), 9), np.float32)
        call("yas_mesher_download", self.h, _ptr(tris), _ptr(normals), keep)
        nb = ctypes.c_int()
        call("yas_mesher_blocks", self.h, None, None, None, None, 0, ctypes.byref(nb))
        nbk = nb.value
        keys = np.zeros((max(nbk, 1), 3), np.int32)
        first = np.zeros(max(nbk, 1), np.int32)
        count = np.zeros(max(nbk, 1), np.int32)
        cases = np.zeros((max(nbk, 1), 512), np.int16) if with_cases else None
        call("yas_mesher_blocks", self.h, _ptr(keys), _ptr(first), _ptr(count),
             _ptr(cases) if with_cases else None, nbk, ctypes.byref(nb))
        out = dict(num_triangles=ntri, tris=tris[:keep], normals=normals[:keep], keys=keys[:nbk],
                   first=first[:nbk], count=count[:nbk])
        if with_cases:
            out["cases"] = cases[:nbk]
        return out


    def WritePLY(self, path):
        """Write the last indexed extraction as binary PLY."""
        call("yas_mesher_write_ply", self.h, str(path).encode())

    def MeshifyIndexed(self, map_struct):
        """Indexed marching cubes: float32 (V, 3) vertices, int32 (T, 3)
        triangles, in the same triangle order as Meshify."""
        nv, nt = ctypes.c_int(), ctypes.c_int()
        call("yas_mesh_extract_indexed", self.h, map_struct.h, ctypes.byref(nv), ctypes.byref(nt), allow=(6,))
        kv, kt = min(nv.value, 3 * self.max_triangles), min(nt.value, self.max_triangles)
        verts = np.zeros((max(kv, 1), 3), np.float32)
        tris = np.zeros((max(kt, 1), 3), np.int32)
        call("yas_mesher_download_indexed", self.h, _ptr(verts), _ptr(tris), kv, kt)
        return dict(num_vertices=nv.value, num_triangles=nt.value, vertices=verts[:kv], triangles=tris[:kt])


def canonical_mesh(m):
    """Reorder a mesh's blocks by key (x, y, z); triangles keep their
    within-block order.  Returns (keys, counts, cases|None, tris, normals)."""
    keys = m["keys"]
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    idx = np.concatenate([np.arange(m["first"][i], m["first"][i] + m["count"][i]) for i in order]) \
        if len(order) else np.zeros(0, np.int64)
    idx = idx.astype(np.int64)
    cases = m["cases"][order] if "cases" in m else None
    return keys[order], m["count"][order], cases, m["tris"][idx], m["normals"][idx]


class RGBDTracking(_Native):
    """RGBDTracking (RGBDFusion/include/RGBDTracking.h:9-107)."""
    _DESTROY = "yas_tracker_destroy"
    _RANK = 2

    def __init__(self, w, h, K, levels=3, iters=(10, 5, 3), mode=TRACK_DEPTH, weights=(1.0, 1.0, 100.0, 10.0),
                 ctx=None):
        self.ctx = ctx or default_context()
        self.w, self.h_ = w, h
        it = (ctypes.c_int * 5)(*(list(iters) + [0] * (5 - len(iters))))
        self.params = TrackerParams(levels, it, mode, *weights)
        Kd = _f64(K)
        h_ = ctypes.c_void_p()
        call("yas_tracker_create", self.ctx.h, w, h, _ptr(Kd), ctypes.byref(self.params), ctypes.byref(h_))
        self._own(h_)

    def SetReferenceImage(self, gray):
        call("yas_tracker_set_reference_image", self.h, _dev(gray), _pitch(gray))

    def SetTrackingImage(self, gray):
        call("yas_tracker_set_tracking_image", self.h, _dev(gray), _pitch(gray))

    def SetReferenceDepth(self, depth):
        call("yas_tracker_set_reference_depth", self.h, _dev(depth), _pitch(depth))

    def SetTrackingDepth(self, depth):
        call("yas_tracker_set_tracking_depth", self.h, _dev(depth), _pitch(depth))

    def SetReferenceModel(self, vmap):
        """vmap: torch (h, w, 4) float32 on the device, or (ptr, pitch)."""
        if isinstance(vmap, tuple):
            call("yas_tracker_set_reference_model", self.h, ctypes.c_void_p(vmap[0]), vmap[1])
        else:
            call("yas_tracker_set_reference_model", self.h, _dev(vmap), _pitch(vmap))

    def GetTransform(self, init=None, swap=True):
        init = np.eye(4) if init is None else init
        out = np.zeros((4, 4))
        st = TrackStats()
        call("yas_tracker_get_transform", self.h, _ptr(_f64(init)), 1 if swap else 0, _ptr(out), ctypes.byref(st),
             allow=(5,))
        return out, st

    def SwapFrameBuffer(self):
        call("yas_tracker_swap_frame_buffer", self.h)

    def IsTrackingGood(self):
        g = ctypes.c_int()
        call("yas_tracker_is_tracking_good", self.h, ctypes.byref(g))
        return bool(g.value)

    def step_sums(self, level, T, kind):
        out = np.zeros(29, np.float32)
        stats = np.zeros(4, np.float32)
        call("yas_tracker_step_sums", self.h, level, _ptr(_f64(T)), kind, _ptr(out), _ptr(stats))
        return out, stats

    def level(self, which, lvl):
        self.ctx.synchronize()
        return _tracker_level(self.h, which, lvl)


def _tracker_level(h_, which, lvl):
    """Host copy of one tracker buffer (which: 0 current intensity, 1
    reference intensity, 2/3 gradients, 4/5 current/reference inverse depth,
    6-9 vertex/normal maps, 10 photometric rows) at pyramid level lvl."""
    p = ctypes.c_void_p()
    pitch = ctypes.c_size_t()
    w, h = ctypes.c_int(), ctypes.c_int()
    call("yas_tracker_level", h_, which, lvl, ctypes.byref(p), ctypes.byref(pitch), ctypes.byref(w),
         ctypes.byref(h))
    if which < 6:
        return download(p.value, (h.value, w.value), np.float32)
    return download(p.value, (h.value, w.value, 4 if which < 10 else 3), np.float32)


class DensePipeline(_Native):
    """Tracking::TrackRGBD's dense steps on one volume (yas_pipeline_*)."""
    _DESTROY = "yas_pipeline_destroy"
    _RANK = 1

    def __init__(self, w, h, K, map_params, tracker_params, depth_scale=1000.0, gray_mode=0, ctx=None):
        self.ctx = ctx or default_context()
        self.w, self.h_ = w, h
        Kf = np.asarray(K, np.float32)
        self.map_params = map_params
        self.tracker_params = tracker_params
        h_ = ctypes.c_void_p()
        call("yas_pipeline_create", self.ctx.h, w, h, _ptr(Kf), ctypes.byref(map_params),
             ctypes.byref(tracker_params), depth_scale, gray_mode, ctypes.byref(h_))
        self._own(h_)

    def step(self, depth_u16, rgb_u8):
        call("yas_pipeline_step", self.h, _dev(depth_u16), _pitch(depth_u16), _dev(rgb_u8), _pitch(rgb_u8))

    def pose(self, frame):
        T = np.zeros((4, 4))
        tr = ctypes.c_int()
        call("yas_pipeline_pose", self.h, frame, _ptr(T), ctypes.byref(tr))
        return T, bool(tr.value)

    def num_frames(self):
        n = ctypes.c_int()
        call("yas_pipeline_num_frames", self.h, ctypes.byref(n))
        return n.value

    def map_handle(self):
        m = ctypes.c_void_p()
        call("yas_pipeline_map", self.h, ctypes.byref(m))
        return m

    def tracker_level(self, which, lvl):
        """A buffer of the pipeline's tracker (see RGBDTracking.level)."""
        t = ctypes.c_void_p()
        call("yas_pipeline_tracker", self.h, ctypes.byref(t))
        self.ctx.synchronize()
        return _tracker_level(t, which, lvl)

    def map_stats(self):
        with _MapView(self.map_handle(), VOXEL_F32_F32, self) as view:
            return view.stats()

    def export_map_sorted(self, voxel_format):
        """Sorted blocks of the pipeline's volume (see MapStruct.export_sorted)."""
        with _MapView(self.map_handle(), voxel_format, self) as view:
            return view.export_sorted()


def map_params(num_buckets, num_blocks, voxel_size, trunc_dist, voxel_format=VOXEL_F32_F32, max_weight=255,
               depth_min=0.1, depth_max=3.0):
    return MapParams(num_buckets, num_blocks, voxel_size, trunc_dist, voxel_format, max_weight, depth_min, depth_max)


def tracker_params(levels=3, iters=(10, 5, 3), mode=TRACK_DEPTH, weights=(1.0, 1.0, 100.0, 10.0)):
    it = (ctypes.c_int * 5)(*(list(iters) + [0] * (5 - len(iters))))
    return TrackerParams(levels, it, mode, *weights)
