"""ctypes bindings for the two C-ABI libraries (include/rtq.h, include/rtq_build.h).

`librtq_build.so` is host-only (builder restatement and synthetic workloads); `librtq.so` is the
gfx950 HIP library (upload, dispatch, read-back, shadow masks, light culling). Both are built in
tree by `tim_raytracing_amd.build`. Loading raises if a library is missing: there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# RTQ_BUILD_LIB names another build of the host library (the tests' sanitizer builds)
BUILD_LIB = os.environ.get("RTQ_BUILD_LIB") or os.path.join(PKG_DIR, "librtq_build.so")
# RTQ_DEVICE_LIB names another build of the device library (a build with other compile-time settings)
DEVICE_LIB = os.environ.get("RTQ_DEVICE_LIB") or os.path.join(PKG_DIR, "librtq.so")

RTQ_OK = 0
STATUS_NAMES = {
    0: "RTQ_OK", 1: "RTQ_ERR_INVALID_ARGUMENT", 2: "RTQ_ERR_HIP", 3: "RTQ_ERR_OUT_OF_MEMORY",
    4: "RTQ_ERR_LIMIT", 5: "RTQ_ERR_BAD_BLOB", 6: "RTQ_ERR_NO_DEVICE", 7: "RTQ_ERR_BUDGET",
    8: "RTQ_ERR_QUEUE_FULL", 9: "RTQ_ERR_CANCELLED", 10: "RTQ_ERR_TIMEOUT", 11: "RTQ_ERR_INTERNAL",
    12: "RTQ_ERR_CACHE_MISS",
}

RTQ_TRACE_CLOSEST = 1
RTQ_TRACE_ANY = 2
RTQ_SHADOW_SUN = 1
RTQ_CAMERA_REFERENCE_ROWS = 1
RTQ_OPTION_PERSISTENT = 1
RTQ_OPTION_BLOCKS_PER_CU = 2
RTQ_OPTION_SORT_RAYS = 3
RTQ_OPTION_LEAF_WAIT = 4
RTQ_NO_HIT = 0xFFFFFFFF
RTQ_PROBE_RAYS = 64
RTQ_BATCH_REFILL = 1
RTQ_ERR_LIMIT = 4
RTQ_ERR_CACHE_MISS = 12
RTQ_SCENE_BOX, RTQ_SCENE_SPHERES, RTQ_SCENE_SOUP = 1, 2, 3


class Range(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("size", C.c_uint32)]


class Ranges(C.Structure):
    _fields_ = [(n, Range) for n in ("triangles", "primitives", "materials", "lights", "nodes", "leaf_data", "blas")]


class Frustum(C.Structure):
    _fields_ = [("camera_pos", C.c_float * 3), ("corner00", C.c_float * 3),
                ("corner10", C.c_float * 3), ("corner01", C.c_float * 3)]


class Lpf(C.Structure):
    _fields_ = [("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3), ("resolution", C.c_uint32 * 3)]


class BuildParams(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("min_obj_per_node", C.c_uint32), ("min_obj_gain", C.c_uint32),
                ("expand_node_volume_threshold", C.c_float), ("expand_node_dimension_factor", C.c_float)]


class BuildStats(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_leaves", C.c_uint32), ("max_leaf_triangles", C.c_uint32),
                ("max_depth", C.c_uint32), ("leaf_triangle_slots", C.c_uint64), ("num_duplicated", C.c_uint32),
                ("num_triangles", C.c_uint32), ("build_seconds", C.c_double)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("flags", C.c_uint32), ("num_rays", C.c_uint32), ("reserved", C.c_uint32),
                ("d_rays", C.c_void_p), ("d_out", C.c_void_p), ("d_leaf_nid", C.c_void_p)]


class BatchParams(C.Structure):
    _fields_ = [("chunk_rays", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("stack_depth", C.c_uint32),
                ("use_graph", C.c_uint32), ("leaf_wait", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_queue", C.c_uint32), ("max_batch_rays", C.c_uint32),
                ("plain", C.c_uint32), ("max_inflight", C.c_uint32), ("params", BatchParams)]


class RequestStats(C.Structure):
    _fields_ = [("queued_ms", C.c_double), ("latency_ms", C.c_double), ("batch_id", C.c_uint32),
                ("batch_requests", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_inner", C.c_uint32), ("num_leaf_slots", C.c_uint32),
                ("num_triangles", C.c_uint32), ("num_lights", C.c_uint32), ("max_depth", C.c_uint32),
                ("device_bytes", C.c_uint64)]


# numpy views of rtq_ray / rtq_hit (32 B and 16 B records)
RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmax", "<f4"), ("dir", "<f4", 3), ("pad", "<u4")])
HIT_DTYPE = np.dtype([("tri_id", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16

P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64
F = C.c_float
FP = C.POINTER(C.c_float)

_BUILD_SIGS = {
    "rtq_build_last_error": ([], C.c_char_p),
    "rtq_builder_create": ([C.POINTER(P)], C.c_int),
    "rtq_builder_destroy": ([P], C.c_int),
    "rtq_builder_add_vertices": ([P, P, U32, C.POINTER(U32)], C.c_int),
    "rtq_builder_add_triangle_list": ([P, U32, U32, P, P], C.c_int),
    "rtq_builder_add_sphere_light": ([P, P, F, P, F], C.c_int),
    "rtq_builder_set_parameters": ([P, C.POINTER(BuildParams)], C.c_int),
    "rtq_builder_set_blas_parameters": ([P, C.POINTER(BuildParams)], C.c_int),
    "rtq_builder_add_blas": ([P, U32, U32, P, P, C.POINTER(U32)], C.c_int),
    "rtq_builder_build": ([P, C.c_int], C.c_int),
    "rtq_builder_gpu_size": ([P, C.POINTER(C.c_size_t)], C.c_int),
    "rtq_builder_fill_gpu_buffer": ([P, P, C.c_size_t, C.POINTER(Ranges)], C.c_int),
    "rtq_builder_positions": ([P, C.POINTER(P), C.POINTER(U32)], C.c_int),
    "rtq_builder_get_stats": ([P, C.POINTER(BuildStats)], C.c_int),
    "rtq_builder_scene_box": ([P, P, P], C.c_int),
    "rtq_scene_generate": ([P, U32, U64, U32], C.c_int),
    "rtq_lights_generate": ([P, U64, U32, F, F, F, F], C.c_int),
    "rtq_camera_frustum": ([P, P, P, F, F, C.POINTER(Frustum)], C.c_int),
    "rtq_frustum_rays": ([C.POINTER(Frustum), U32, U32, U64, F, P], C.c_int),
    "rtq_surface_rays": ([P, U64, U32, F, P], C.c_int),
    "rtq_bounce_rays": ([P, U64, P, P, U32, F, P], C.c_int),
    "rtq_probe_directions": ([U64, U32, P], C.c_int),
    "rtq_builder_input_hash": ([P, C.POINTER(U64)], C.c_int),
    "rtq_cache_save": ([P, C.c_char_p], C.c_int),
    "rtq_cache_load": ([C.c_char_p, U64, C.POINTER(P)], C.c_int),
    "rtq_cached_scene_view": ([P, C.POINTER(P), C.POINTER(C.c_size_t), C.POINTER(Ranges), C.POINTER(P),
                               C.POINTER(U32), C.POINTER(U64)], C.c_int),
    "rtq_cached_scene_free": ([P], C.c_int),
}

_DEVICE_SIGS = {
    "rtq_status_string": ([C.c_int], C.c_char_p),
    "rtq_last_error": ([], C.c_char_p),
    "rtq_device_create": ([C.c_int, C.POINTER(P)], C.c_int),
    "rtq_device_destroy": ([P], C.c_int),
    "rtq_device_alloc": ([P, C.c_size_t, C.POINTER(P)], C.c_int),
    "rtq_device_free": ([P, P], C.c_int),
    "rtq_host_alloc": ([P, C.c_size_t, C.POINTER(P)], C.c_int),
    "rtq_host_free": ([P, P], C.c_int),
    "rtq_copy_to_device": ([P, P, P, C.c_size_t, P], C.c_int),
    "rtq_device_copy": ([P, P, P, C.c_size_t, P], C.c_int),
    "rtq_device_set_option": ([P, U32, U32], C.c_int),
    "rtq_device_get_option": ([P, U32, C.POINTER(U32)], C.c_int),
    "rtq_stream_create": ([P, C.POINTER(P)], C.c_int),
    "rtq_stream_destroy": ([P, P], C.c_int),
    "rtq_stream_synchronize": ([P, P], C.c_int),
    "rtq_event_create": ([P, C.POINTER(P)], C.c_int),
    "rtq_event_destroy": ([P, P], C.c_int),
    "rtq_event_record": ([P, P], C.c_int),
    "rtq_event_elapsed_ms": ([P, P, C.POINTER(C.c_float)], C.c_int),
    "rtq_scene_upload": ([P, P, C.c_size_t, C.POINTER(Ranges), P, U32, C.POINTER(P)], C.c_int),
    "rtq_scene_destroy": ([P], C.c_int),
    "rtq_scene_get_info": ([P, C.POINTER(SceneInfo)], C.c_int),
    "rtq_dispatch_rays": ([P, P, U32, U32, P, P, P], C.c_int),
    "rtq_dispatch_rays_counted": ([P, P, U32, U32, P, P, P, P], C.c_int),
    "rtq_read_hits": ([P, P, P, C.c_size_t, P], C.c_int),
    "rtq_trace_camera": ([P, C.POINTER(Frustum), U32, U32, U32, P, P, P], C.c_int),
    "rtq_trace_camera_record": ([P, C.POINTER(Frustum), U32, U32, U32, P, U32, C.POINTER(Lpf), P, P], C.c_int),
    "rtq_trace_probe_field": ([P, C.POINTER(Lpf), P, P, U32, P, P], C.c_int),
    "rtq_shadow_masks": ([P, P, P, P, U32, P, U32, P, P], C.c_int),
    "rtq_cull_lights": ([P, C.POINTER(Frustum), U32, U32, P, P, U32, C.POINTER(U32), P], C.c_int),
    "rtq_cull_lights_noalloc": ([P, C.POINTER(Frustum), U32, U32, P, P, U32, P], C.c_int),
    "rtq_selftest_division": ([P, U64, U64, C.POINTER(U64)], C.c_int),
    "rtq_batch_create": ([P, P, U32, P, U32, C.POINTER(P)], C.c_int),
    "rtq_batch_destroy": ([P], C.c_int),
    "rtq_batch_set_params": ([P, C.POINTER(BatchParams)], C.c_int),
    "rtq_batch_get_params": ([P, C.POINTER(BatchParams)], C.c_int),
    "rtq_batch_launch": ([P, P], C.c_int),
    "rtq_batch_launch_plain": ([P, P], C.c_int),
    "rtq_batch_autotune": ([P, U32, C.POINTER(BatchParams), C.POINTER(C.c_float), P], C.c_int),
    "rtq_server_create": ([P, P, U32, C.POINTER(ServerConfig), C.POINTER(P)], C.c_int),
    "rtq_server_destroy": ([P], C.c_int),
    "rtq_server_submit": ([P, U32, U32, P, U32, P, P, C.POINTER(P)], C.c_int),
    "rtq_server_submit_device": ([P, U32, U32, P, U32, P, P, C.POINTER(P)], C.c_int),
    "rtq_request_wait": ([P, U32], C.c_int),
    "rtq_request_cancel": ([P], C.c_int),
    "rtq_request_get_stats": ([P, C.POINTER(RequestStats)], C.c_int),
    "rtq_request_release": ([P], C.c_int),
}


class RtqError(RuntimeError):
    def __init__(self, fn: str, status: int, detail: str = ""):
        self.status = status
        super().__init__(f"{fn} failed: {STATUS_NAMES.get(status, status)} {detail}".strip())


def _bind(lib, sigs):
    for name, (argtypes, restype) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    return lib


_build_lib = None
_device_lib = None


def bind_build_lib(path: str):
    """Another build of the host builder library at `path` (the tests' planted-mistake builds)."""
    return _bind(C.CDLL(path), _BUILD_SIGS)


def build_lib():
    """The host builder library; raises if it has not been built."""
    global _build_lib
    if _build_lib is None:
        if not os.path.exists(BUILD_LIB):
            raise RuntimeError(f"{BUILD_LIB} is missing: run tim_raytracing_amd.build.build_all() first")
        _build_lib = _bind(C.CDLL(BUILD_LIB), _BUILD_SIGS)
    return _build_lib


def device_lib():
    """The gfx950 HIP library; raises if it has not been built (no CPU fallback exists)."""
    global _device_lib
    if _device_lib is None:
        if not os.path.exists(DEVICE_LIB):
            raise RuntimeError(f"{DEVICE_LIB} is missing: run tim_raytracing_amd.build.build_all() first")
        _device_lib = _bind(C.CDLL(DEVICE_LIB), _DEVICE_SIGS)
    return _device_lib


def check_build(fn: str, status: int):
    if status != RTQ_OK:
        raise RtqError(fn, status, build_lib().rtq_build_last_error().decode())


def check_device(fn: str, status: int):
    if status != RTQ_OK:
        raise RtqError(fn, status, device_lib().rtq_last_error().decode())


def ptr(a) -> int:
    """Address of a contiguous numpy array or torch tensor."""
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    if a is None:
        return None
    return a.data_ptr()
