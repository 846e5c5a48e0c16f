"""ctypes loader of libyas.so (the HIP kernels behind include/yas.h).

The product path has no fallback: if the library is missing or fails to load,
every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyas.so")

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double
SZ = ctypes.c_size_t
PP = ctypes.POINTER(ctypes.c_void_p)
PI = ctypes.POINTER(ctypes.c_int)


class MapParams(ctypes.Structure):
    _fields_ = [("num_buckets", I), ("num_blocks", I), ("voxel_size", F), ("trunc_dist", F),
                ("voxel_format", I), ("max_weight", I), ("depth_min", F), ("depth_max", F)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("alloc_tile_w", I), ("visible_groups", I), ("integrate_threads", I), ("integrate_groups", I),
                ("step_ms", F)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Result(ctypes.Structure):
    _fields_ = [("stream", I), ("tracked", I), ("request", ctypes.c_longlong), ("pose", ctypes.c_double * 16)]


class ServerStats(ctypes.Structure):
    _fields_ = [("streams", I), ("frames_done", ctypes.c_longlong), ("batches_done", ctypes.c_longlong),
                ("bytes_in_use", ctypes.c_size_t), ("bytes_budget", ctypes.c_size_t), ("streams_pending", I),
                ("streams_hibernated", I), ("hibernations", ctypes.c_longlong), ("reactivations", ctypes.c_longlong)]


class MapStats(ctypes.Structure):
    _fields_ = [("num_allocated", I), ("num_visible", I), ("flags", I), ("reserved", I)]


class TrackerParams(ctypes.Structure):
    _fields_ = [("levels", I), ("iters", I * 5), ("mode", I), ("w_rgb_h", F), ("w_rgb_b", F),
                ("w_icp_h", F), ("w_icp_b", F)]


class TrackStats(ctypes.Structure):
    _fields_ = [("good", I), ("icp_inliers", I), ("rgb_inliers", I), ("last_error", F)]


# name -> (restype, argtypes); every symbol declared in include/yas.h.
SIGNATURES = {
    "yas_version": (ctypes.c_char_p, []),
    "yas_last_error": (ctypes.c_char_p, []),
    "yas_ctx_create": (I, [I, P, PP]),
    "yas_ctx_create_null_stream": (I, [I, PP]),
    "yas_ctx_destroy": (I, [P]),
    "yas_ctx_synchronize": (I, [P]),
    "yas_ctx_stream": (P, [P]),
    "yas_map_create": (I, [P, ctypes.POINTER(MapParams), P, PP]),
    "yas_map_reset": (I, [P]),
    "yas_map_destroy": (I, [P]),
    "yas_map_set_pose": (I, [P, P]),
    "yas_map_get_pose": (I, [P, P]),
    "yas_map_fuse": (I, [P, P, SZ, I, I, P]),
    "yas_map_fuse_u16": (I, [P, P, SZ, I, I, F, P]),
    "yas_map_check_visible_blocks": (I, [P, I, I, P, PI]),
    "yas_map_get_stats": (I, [P, ctypes.POINTER(MapStats)]),
    "yas_map_visible_blocks": (I, [P, P, I, PI]),
    "yas_map_export": (I, [P, P, P, I, PI]),
    "yas_map_voxel_bytes": (I, [P]),
    "yas_raycaster_create": (I, [P, I, I, P, PP]),
    "yas_raycaster_destroy": (I, [P]),
    "yas_raycast": (I, [P, P, P]),
    "yas_raycaster_vmap": (I, [P, PP, ctypes.POINTER(SZ)]),
    "yas_raycaster_depth": (I, [P, PP, ctypes.POINTER(SZ), P, SZ]),
    "yas_raycast_into": (I, [P, P, P, I, I, P, SZ]),
    "yas_raycaster_zrange": (I, [P, PP, PP, PI, PI]),
    "yas_raycaster_num_visible": (I, [P, PI]),
    "yas_mesher_create": (I, [P, I, PP]),
    "yas_mesher_destroy": (I, [P]),
    "yas_mesh_extract": (I, [P, P, PI]),
    "yas_mesher_download": (I, [P, P, P, I]),
    "yas_mesh_extract_indexed": (I, [P, P, PI, PI]),
    "yas_mesher_write_ply": (I, [P, ctypes.c_char_p]),
    "yas_map_fuse_map": (I, [P, P]),
    "yas_server_create": (I, [P, I, I, I, P, I, ctypes.c_size_t, P]),
    "yas_server_destroy": (I, [P]),
    "yas_server_add_stream": (I, [P, I, I, P, P, ctypes.c_float, I, PI]),
    "yas_server_open_stream": (I, [P, I, I, P, P, F, I, PI]),
    "yas_server_close_stream": (I, [P, I]),
    "yas_server_set_eviction": (I, [P, D]),
    "yas_server_stream_state": (I, [P, I, PI]),
    "yas_server_remove_stream": (I, [P, I]),
    "yas_server_stream_map": (I, [P, I, P]),
    "yas_server_stream_pipeline": (I, [P, I, P]),
    "yas_server_set_max_batch": (I, [P, I]),
    "yas_server_set_gn_group": (I, [P, I]),
    "yas_server_set_max_in_flight": (I, [P, I, I]),
    "yas_server_prepare": (I, [P]),
    "yas_server_submit": (I, [P, I, P, ctypes.c_size_t, P, ctypes.c_size_t, P]),
    "yas_server_pump": (I, [P, PI]),
    "yas_server_poll": (I, [P, P, I, PI]),
    "yas_server_drain": (I, [P]),
    "yas_server_get_stats": (I, [P, P]),
    "yas_batch_create": (I, [P, I, P, P]),
    "yas_batch_destroy": (I, [P]),
    "yas_batch_plan": (I, [P, I, P, P, P, P, ctypes.c_float, P]),
    "yas_batch_run": (I, [P, I, I]),
    "yas_batch_autotune": (I, [P, I, P]),
    "yas_batch_set_config": (I, [P, P]),
    "yas_batch_get_config": (I, [P, P]),
    "yas_map_get_params": (I, [P, P, P]),
    "yas_map_reserve": (I, [P, I, I]),
    "yas_map_hibernate": (I, [P]),
    "yas_map_reactivate": (I, [P]),
    "yas_map_is_hibernating": (I, [P, PI]),
    "yas_map_save": (I, [P, ctypes.c_char_p]),
    "yas_map_load": (I, [P, ctypes.c_char_p, P]),
    "yas_mesher_download_indexed": (I, [P, P, P, I, I]),
    "yas_mesher_indexed_device": (I, [P, P, P]),
    "yas_mesher_blocks": (I, [P, P, P, P, P, I, PI]),
    "yas_tracker_create": (I, [P, I, I, P, ctypes.POINTER(TrackerParams), PP]),
    "yas_tracker_destroy": (I, [P]),
    "yas_tracker_set_reference_image": (I, [P, P, SZ]),
    "yas_tracker_set_tracking_image": (I, [P, P, SZ]),
    "yas_tracker_set_reference_depth": (I, [P, P, SZ]),
    "yas_tracker_set_tracking_depth": (I, [P, P, SZ]),
    "yas_tracker_set_reference_image_host": (I, [P, P, SZ]),
    "yas_tracker_set_tracking_image_host": (I, [P, P, SZ]),
    "yas_tracker_set_reference_depth_host": (I, [P, P, SZ]),
    "yas_tracker_set_tracking_depth_host": (I, [P, P, SZ]),
    "yas_tracker_raw_depth": (I, [P, PP, ctypes.POINTER(SZ)]),
    "yas_tracker_write_debug_images": (I, [P, ctypes.c_char_p]),
    "yas_tracker_set_reference_model": (I, [P, P, SZ]),
    "yas_tracker_get_transform": (I, [P, P, I, P, ctypes.POINTER(TrackStats)]),
    "yas_tracker_swap_frame_buffer": (I, [P]),
    "yas_tracker_is_tracking_good": (I, [P, PI]),
    "yas_tracker_step_sums": (I, [P, I, P, I, P, P]),
    "yas_tracker_solve6": (I, [P, P, P, P, I]),
    "yas_tracker_level": (I, [P, I, I, PP, ctypes.POINTER(SZ), PI, PI]),
    "yas_depth_u16_to_f32": (I, [P, P, SZ, P, SZ, I, I, F]),
    "yas_rgb_to_intensity": (I, [P, P, SZ, P, SZ, I, I, I]),
    "yas_depth_to_invdepth": (I, [P, P, SZ, P, SZ, I, I]),
    "yas_pyrdown_depth": (I, [P, P, SZ, P, SZ, I, I]),
    "yas_pyrdown_gauss": (I, [P, P, SZ, I, I, P, SZ]),
    "yas_compute_vertex_map": (I, [P, P, SZ, P, SZ, I, I, P, F]),
    "yas_compute_normal_map": (I, [P, P, SZ, P, SZ, I, I]),
    "yas_image_gradient": (I, [P, P, SZ, P, P, SZ, I, I]),
    "yas_vmap_to_depth": (I, [P, P, SZ, P, SZ, I, I]),
    "yas_pipeline_create": (I, [P, I, I, P, ctypes.POINTER(MapParams), ctypes.POINTER(TrackerParams), F, I, PP]),
    "yas_pipeline_destroy": (I, [P]),
    "yas_pipeline_step": (I, [P, P, SZ, P, SZ]),
    "yas_pipeline_pose": (I, [P, I, P, PI]),
    "yas_pipeline_num_frames": (I, [P, PI]),
    "yas_pipeline_enable_timing": (I, [P, I]),
    "yas_pipeline_timing": (I, [P, P, P, I, PI]),
    "yas_pipeline_reset_timing": (I, [P]),
    "yas_pipeline_kernel_timing": (I, [P, P, PI]),
    "yas_pipeline_frame_stats": (I, [P, I, I, P]),
    "yas_pipeline_map": (I, [P, PP]),
    "yas_pipeline_raycaster": (I, [P, PP]),
    "yas_pipeline_tracker": (I, [P, PP]),
}

STATUS = {0: "YAS_OK", 1: "YAS_E_INVALID", 2: "YAS_E_HIP", 3: "YAS_E_OUT_OF_BLOCKS", 4: "YAS_E_HASH_FULL",
          5: "YAS_E_DIVERGED", 6: "YAS_E_CAPACITY", 7: "YAS_E_REJECTED"}


class YasError(RuntimeError):
    def __init__(self, name, code, msg):
        super().__init__("%s -> %s: %s" % (name, STATUS.get(code, code), msg))
        self.code = code


_lib = None


def lib():
    """Load libyas.so; raises if it is not built (there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libyas.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
    # torch bundles a HIP runtime with the same soname as /opt/rocm's; loading
    # torch first makes libyas.so bind to that one runtime instead of a second.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def call(name, *args, allow=()):
    """Call an int-returning entry point and raise YasError on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0 and rc not in allow:
        raise YasError(name, rc, (lib().yas_last_error() or b"").decode())
    return rc
