"""Python view of the gfx950 library (librtq.so) over its C ABI.

Mirrors the reference's upload / dispatch surface: Scene.upload stands for BVHData::build's
CreateBuffer + UploadBuffer (Renderer/BVHData.cpp:43-48), Scene.dispatch for RayTracingPass::tracePass
-> IRenderContext::Dispatch (Renderer/raytracingPass.cpp:89-110). There is no CPU fallback: every
call goes to the HIP library and raises if it is missing or fails.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import bindings as B


class DeviceBuffer:
    def __init__(self, dev: "Device", nbytes: int):
        self.dev = dev
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        B.check_device("rtq_device_alloc", dev.lib.rtq_device_alloc(dev.handle, max(1, self.nbytes), C.byref(p)))
        self.ptr = p.value

    def free(self):
        if self.ptr is not None:
            self.dev.lib.rtq_device_free(self.dev.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def write(self, arr: np.ndarray, stream=None):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        B.check_device("rtq_copy_to_device",
                       self.dev.lib.rtq_copy_to_device(self.dev.handle, self.ptr, a.ctypes.data, a.nbytes, stream))

    def read(self, dtype, count: int, stream=None) -> np.ndarray:
        out = np.zeros(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        B.check_device("rtq_read_hits",
                       self.dev.lib.rtq_read_hits(self.dev.handle, self.ptr, out.ctypes.data, out.nbytes, stream))
        return out


class DeviceView(DeviceBuffer):
    """A byte range of a DeviceBuffer; keeps the buffer alive and frees nothing itself."""

    def __init__(self, parent: DeviceBuffer, offset: int, nbytes: int):
        assert 0 <= offset and offset + nbytes <= parent.nbytes
        self.dev = parent.dev
        self.parent = parent
        self.nbytes = int(nbytes)
        self.ptr = parent.ptr + int(offset)

    def free(self):
        self.ptr = None


class Stream:
    """A HIP stream of the library's runtime (default priority)."""

    def __init__(self, dev: "Device"):
        self.dev = dev
        h = C.c_void_p()
        B.check_device("rtq_stream_create", dev.lib.rtq_stream_create(dev.handle, C.byref(h)))
        self.ptr = h.value

    def synchronize(self):
        B.check_device("rtq_stream_synchronize", self.dev.lib.rtq_stream_synchronize(self.dev.handle, self.ptr))

    def __del__(self):
        try:
            if self.ptr:
                self.dev.lib.rtq_stream_destroy(self.dev.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass


class Event:
    def __init__(self, dev: "Device"):
        self.dev = dev
        h = C.c_void_p()
        B.check_device("rtq_event_create", dev.lib.rtq_event_create(dev.handle, C.byref(h)))
        self.ptr = h.value

    def record(self, stream=None):
        B.check_device("rtq_event_record", self.dev.lib.rtq_event_record(self.ptr, stream))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        B.check_device("rtq_event_elapsed_ms", self.dev.lib.rtq_event_elapsed_ms(self.ptr, end.ptr, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.ptr:
                self.dev.lib.rtq_event_destroy(self.dev.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass


class HostBuffer:
    """Page-locked host memory from the library's runtime, viewed as a numpy array."""

    def __init__(self, dev: "Device", nbytes: int):
        self.dev = dev
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        B.check_device("rtq_host_alloc", dev.lib.rtq_host_alloc(dev.handle, max(1, self.nbytes), C.byref(p)))
        self.ptr = p.value

    def view(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        dt = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dt.itemsize
        assert offset + count * dt.itemsize <= self.nbytes
        raw = (C.c_uint8 * (count * dt.itemsize)).from_address(self.ptr + offset)
        return np.frombuffer(raw, dtype=dt, count=count)

    def free(self):
        if self.ptr is not None:
            self.dev.lib.rtq_host_free(self.dev.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Device:
    def __init__(self, ordinal: int = 0):
        self.lib = B.device_lib()
        h = C.c_void_p()
        B.check_device("rtq_device_create", self.lib.rtq_device_create(ordinal, C.byref(h)))
        self.handle = h

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def host_alloc(self, nbytes: int) -> HostBuffer:
        return HostBuffer(self, nbytes)

    def stream(self) -> Stream:
        return Stream(self)

    def set_option(self, option: int, value: int):
        B.check_device("rtq_device_set_option", self.lib.rtq_device_set_option(self.handle, option, value))

    def get_option(self, option: int) -> int:
        v = C.c_uint32()
        B.check_device("rtq_device_get_option", self.lib.rtq_device_get_option(self.handle, option, C.byref(v)))
        return v.value

    def event(self) -> Event:
        return Event(self)

    def upload(self, arr: np.ndarray) -> DeviceBuffer:
        buf = DeviceBuffer(self, np.asarray(arr).nbytes)
        buf.write(arr)
        return buf

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.rtq_device_destroy(h)
            self.handle = None


class Batch:
    """Batched form: many (scene, ray set) items in one persistent launch (rtq_batch_*)."""

    def __init__(self, dev: Device, scenes, items):
        """items: list of (scene_index, flags, d_rays_ptr, num_rays, d_out_ptr, d_nid_ptr_or_None)"""
        self.dev = dev
        arr = (C.c_void_p * max(1, len(scenes)))(*[s.handle.value for s in scenes])
        its = (B.BatchItem * max(1, len(items)))()
        for k, (si, flags, d_rays, n, d_out, d_nid) in enumerate(items):
            its[k] = B.BatchItem(si, flags, n, 0, d_rays, d_out, d_nid)
        h = C.c_void_p()
        B.check_device("rtq_batch_create", dev.lib.rtq_batch_create(dev.handle, arr, len(scenes), its, len(items),
                                                                    C.byref(h)))
        self.handle = h
        self._scenes = list(scenes)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.dev.lib.rtq_batch_destroy(h)
            self.handle = None

    def params(self) -> B.BatchParams:
        p = B.BatchParams()
        B.check_device("rtq_batch_get_params", self.dev.lib.rtq_batch_get_params(self.handle, C.byref(p)))
        return p

    def set_params(self, chunk_rays=None, blocks_per_cu=None, stack_depth=None, use_graph=None, leaf_wait=None):
        p = self.params()
        if chunk_rays is not None:
            p.chunk_rays = chunk_rays
        if blocks_per_cu is not None:
            p.blocks_per_cu = blocks_per_cu
        if stack_depth is not None:
            p.stack_depth = stack_depth
        if use_graph is not None:
            p.use_graph = int(use_graph)
        if leaf_wait is not None:
            p.leaf_wait = leaf_wait
        B.check_device("rtq_batch_set_params", self.dev.lib.rtq_batch_set_params(self.handle, C.byref(p)))

    def launch(self, stream=None):
        B.check_device("rtq_batch_launch", self.dev.lib.rtq_batch_launch(self.handle, stream))

    def launch_plain(self, stream=None):
        B.check_device("rtq_batch_launch_plain", self.dev.lib.rtq_batch_launch_plain(self.handle, stream))

    def autotune(self, reps=2, stream=None):
        p = B.BatchParams()
        ms = C.c_float()
        B.check_device("rtq_batch_autotune", self.dev.lib.rtq_batch_autotune(self.handle, reps, C.byref(p),
                                                                             C.byref(ms), stream))
        return p, ms.value


def check_request_buffers(flags: int, rays: np.ndarray, out: np.ndarray, nid: np.ndarray | None) -> int:
    """Validates a served request's host buffers before their addresses reach the library, which reads
    32 B per ray and writes 16 B (closest) or 4 B (any) per ray and 4 B per leaf id: rays must be a
    contiguous RAY_DTYPE array, out a contiguous HIT_DTYPE (closest) or uint32 (any) array and nid, if
    given, a contiguous uint32 array, both of at least as many entries as there are rays. Returns the
    ray count; raises ValueError otherwise."""
    if flags not in (B.RTQ_TRACE_CLOSEST, B.RTQ_TRACE_ANY):
        raise ValueError(f"bad flags {flags}")
    if not isinstance(rays, np.ndarray) or rays.dtype != B.RAY_DTYPE or rays.ndim != 1 or \
            not rays.flags["C_CONTIGUOUS"]:
        raise ValueError("rays must be a contiguous 1-D array of RAY_DTYPE records")
    n = rays.size
    if n == 0:
        raise ValueError("empty request")
    want = B.HIT_DTYPE if flags == B.RTQ_TRACE_CLOSEST else np.dtype(np.uint32)
    if not isinstance(out, np.ndarray) or out.dtype != want or not out.flags["C_CONTIGUOUS"] or \
            not out.flags["WRITEABLE"] or out.size < n:
        raise ValueError(f"out must be a writeable contiguous {want} array of at least {n} entries")
    if nid is not None:
        if flags != B.RTQ_TRACE_CLOSEST:
            raise ValueError("leaf node ids exist for closest-hit requests only")
        if not isinstance(nid, np.ndarray) or nid.dtype != np.uint32 or not nid.flags["C_CONTIGUOUS"] or \
                not nid.flags["WRITEABLE"] or nid.size < n:
            raise ValueError(f"nid must be a writeable contiguous uint32 array of at least {n} entries")
    return n


class Request:
    """A submitted request. The library reads the rays and writes the results asynchronously, so the
    request keeps its host buffers alive; a request dropped before it finishes is cancelled if it is
    still queued, otherwise waited for, and released before those buffers can go."""

    def __init__(self, server: "Server", handle, rays, out, nid):
        self.server = server
        self.handle = handle
        self._keep = (rays, out, nid)  # host buffers must outlive the request
        self.out = out
        self.nid = nid

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib = self.server.dev.lib
                if lib.rtq_request_cancel(self.handle) != B.RTQ_OK:
                    lib.rtq_request_wait(self.handle, 0xFFFFFFFF)
                lib.rtq_request_release(self.handle)
                self.handle = None
        except Exception:
            pass

    def wait(self, timeout_ms: int = 0xFFFFFFFF) -> int:
        return self.server.dev.lib.rtq_request_wait(self.handle, timeout_ms)

    def cancel(self) -> int:
        return self.server.dev.lib.rtq_request_cancel(self.handle)

    def stats(self) -> B.RequestStats:
        st = B.RequestStats()
        B.check_device("rtq_request_get_stats", self.server.dev.lib.rtq_request_get_stats(self.handle, C.byref(st)))
        return st

    def release(self):
        if self.handle:
            B.check_device("rtq_request_release", self.server.dev.lib.rtq_request_release(self.handle))
            self.handle = None


class Server:
    """Served form (rtq_server_*): resident scenes, HBM-budget admission, continuous batching."""

    def __init__(self, dev: Device, scenes, budget_bytes: int, max_queue: int = 0, max_batch_rays: int = 0,
                 plain: bool = False, params: B.BatchParams | None = None, max_inflight: int = 0):
        self.dev = dev
        arr = (C.c_void_p * max(1, len(scenes)))(*[s.handle.value for s in scenes])
        cfg = B.ServerConfig(budget_bytes, max_queue, max_batch_rays, int(plain), max_inflight,
                             params or B.BatchParams())
        h = C.c_void_p()
        B.check_device("rtq_server_create", dev.lib.rtq_server_create(dev.handle, arr, len(scenes), C.byref(cfg),
                                                                      C.byref(h)))
        self.handle = h
        self._scenes = list(scenes)

    def submit(self, scene: int, flags: int, rays: np.ndarray, out: np.ndarray | None = None,
               nid: np.ndarray | None = None, with_nid: bool = True) -> Request:
        n = rays.size
        if out is None:
            out = np.zeros(n, B.HIT_DTYPE) if flags == B.RTQ_TRACE_CLOSEST else np.zeros(n, np.uint32)
        if nid is None and flags == B.RTQ_TRACE_CLOSEST and with_nid:
            nid = np.zeros(n, np.uint32)
        n = check_request_buffers(flags, rays, out, nid)
        h = C.c_void_p()
        B.check_device("rtq_server_submit",
                       self.dev.lib.rtq_server_submit(self.handle, scene, flags, rays.ctypes.data, n, out.ctypes.data,
                                                      nid.ctypes.data if nid is not None else None, C.byref(h)))
        return Request(self, h, rays, out, nid)

    def submit_device(self, scene: int, flags: int, d_rays, n: int, d_out, d_nid=None) -> Request:
        """A request whose rays and results are device buffers (DeviceBuffer objects, kept alive by the
        request): traced in place, nothing crosses the host link."""
        need = 16 * n if flags == B.RTQ_TRACE_CLOSEST else 4 * n
        if d_rays.nbytes < 32 * n or d_out.nbytes < need or (d_nid is not None and d_nid.nbytes < 4 * n):
            raise ValueError("device buffers smaller than the request")
        h = C.c_void_p()
        B.check_device("rtq_server_submit_device",
                       self.dev.lib.rtq_server_submit_device(self.handle, scene, flags, d_rays.ptr, n, d_out.ptr,
                                                             d_nid.ptr if d_nid is not None else None, C.byref(h)))
        return Request(self, h, d_rays, d_out, d_nid)

    def close(self):
        if getattr(self, "handle", None):
            self.dev.lib.rtq_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    def __init__(self, dev: Device, blob: np.ndarray, ranges, positions: np.ndarray):
        self.dev = dev
        blob = np.ascontiguousarray(blob, np.uint8)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        h = C.c_void_p()
        B.check_device("rtq_scene_upload",
                       dev.lib.rtq_scene_upload(dev.handle, blob.ctypes.data, blob.size, C.byref(ranges),
                                                pos.ctypes.data, pos.size // 3, C.byref(h)))
        self.handle = h

    @classmethod
    def from_builder(cls, dev: Device, builder) -> "Scene":
        blob, ranges = builder.fill_gpu_buffer()
        return cls(dev, blob, ranges, builder.positions())

    def info(self) -> B.SceneInfo:
        i = B.SceneInfo()
        B.check_device("rtq_scene_get_info", self.dev.lib.rtq_scene_get_info(self.handle, C.byref(i)))
        return i

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.dev.lib.rtq_scene_destroy(h)
            self.handle = None

    # dispatch-rays on device buffers (asynchronous on `stream`)
    def dispatch(self, d_rays: int, n: int, flags: int, d_out: int, d_nid: int | None = None, stream=None):
        B.check_device("rtq_dispatch_rays",
                       self.dev.lib.rtq_dispatch_rays(self.handle, d_rays, n, flags, d_out, d_nid, stream))

    def dispatch_counted(self, d_rays: int, n: int, flags: int, d_out: int, d_nid, d_counters: int, stream=None):
        B.check_device("rtq_dispatch_rays_counted",
                       self.dev.lib.rtq_dispatch_rays_counted(self.handle, d_rays, n, flags, d_out, d_nid,
                                                              d_counters, stream))

    def shadow_masks(self, d_rays, d_hits, d_nid, n, sun_dir, flags, d_masks, stream=None):
        sd = np.asarray(sun_dir if sun_dir is not None else (0, 0, 0), np.float32)
        B.check_device("rtq_shadow_masks",
                       self.dev.lib.rtq_shadow_masks(self.handle, d_rays, d_hits, d_nid, n, sd.ctypes.data, flags,
                                                     d_masks, stream))

    # ---- host-array conveniences (allocate, copy, run, read back) ----
    def trace_closest(self, rays: np.ndarray, counted: bool = False):
        n = rays.size
        d_r = self.dev.upload(rays)
        d_h = self.dev.alloc(16 * max(n, 1))
        d_n = self.dev.alloc(4 * max(n, 1))
        out = {}
        if counted:
            d_c = self.dev.alloc(12 * max(n, 1))
            self.dispatch_counted(d_r.ptr, n, B.RTQ_TRACE_CLOSEST, d_h.ptr, d_n.ptr, d_c.ptr)
            out["counters"] = d_c.read(np.uint32, 3 * n).reshape(n, 3)
        else:
            self.dispatch(d_r.ptr, n, B.RTQ_TRACE_CLOSEST, d_h.ptr, d_n.ptr)
        hits = d_h.read(B.HIT_DTYPE, n)
        out.update(tri_id=hits["tri_id"].copy(), t=hits["t"].copy(), u=hits["u"].copy(), v=hits["v"].copy(),
                   nid=d_n.read(np.uint32, n), hits=hits)
        return out

    def trace_camera(self, frustum: B.Frustum, width: int, height: int, flags: int = 0):
        n = width * height
        d_h = self.dev.alloc(16 * n)
        d_n = self.dev.alloc(4 * n)
        zeros = np.zeros(n, B.HIT_DTYPE)
        zeros["tri_id"] = 0xFFFFFFFF
        d_h.write(zeros)
        B.check_device("rtq_trace_camera",
                       self.dev.lib.rtq_trace_camera(self.handle, C.byref(frustum), width, height, flags, d_h.ptr,
                                                     d_n.ptr, None))
        hits = d_h.read(B.HIT_DTYPE, n)
        return {"tri_id": hits["tri_id"].copy(), "t": hits["t"].copy(), "u": hits["u"].copy(),
                "v": hits["v"].copy(), "nid": d_n.read(np.uint32, n), "hits": hits}

    def trace_record(self, frustum: B.Frustum, width: int, height: int, flags: int = 0, sun_dir=None,
                     shadow_flags: int = 0, lpf=None) -> np.ndarray:
        """The camera pass's TRACING_STEP record, (H*W, 8) uint32. lpf = (min3, max3, res3) or None.
        Pixels the pass does not write (row 0 with reference rows) read back as 0xFFFFFFFF."""
        n = width * height
        d_r = self.dev.alloc(32 * max(n, 1))
        d_r.write(np.full(8 * n, 0xFFFFFFFF, np.uint32))
        sd = np.asarray(sun_dir if sun_dir is not None else (0, 0, 0), np.float32)
        lp = None
        if lpf is not None:
            lp = B.Lpf((C.c_float * 3)(*lpf[0]), (C.c_float * 3)(*lpf[1]), (C.c_uint32 * 3)(*lpf[2]))
        B.check_device("rtq_trace_camera_record",
                       self.dev.lib.rtq_trace_camera_record(self.handle, C.byref(frustum), width, height, flags,
                                                            sd.ctypes.data if sun_dir is not None else None,
                                                            shadow_flags, C.byref(lp) if lp is not None else None,
                                                            d_r.ptr, None))
        return d_r.read(np.uint32, 8 * n).reshape(n, 8)

    def trace_probe_field(self, lpf, dirs: np.ndarray, sun_dir=None, shadow_flags: int = 0) -> np.ndarray:
        """The light-probe field's TRACING_STEP records, (probes*64, 8) uint32; lpf = (min3, max3, res3)."""
        n = int(np.prod(lpf[2])) * B.RTQ_PROBE_RAYS
        d_r = self.dev.alloc(32 * max(n, 1))
        lp = B.Lpf((C.c_float * 3)(*lpf[0]), (C.c_float * 3)(*lpf[1]), (C.c_uint32 * 3)(*lpf[2]))
        d = np.ascontiguousarray(dirs, np.float32).reshape(B.RTQ_PROBE_RAYS, 4)
        sd = np.asarray(sun_dir if sun_dir is not None else (0, 0, 0), np.float32)
        B.check_device("rtq_trace_probe_field",
                       self.dev.lib.rtq_trace_probe_field(self.handle, C.byref(lp), d.ctypes.data,
                                                          sd.ctypes.data if sun_dir is not None else None,
                                                          shadow_flags, d_r.ptr, None))
        return d_r.read(np.uint32, 8 * n).reshape(n, 8)

    def trace_any(self, rays: np.ndarray, counted: bool = False):
        n = rays.size
        d_r = self.dev.upload(rays)
        d_o = self.dev.alloc(4 * max(n, 1))
        if counted:
            d_c = self.dev.alloc(12 * max(n, 1))
            self.dispatch_counted(d_r.ptr, n, B.RTQ_TRACE_ANY, d_o.ptr, None, d_c.ptr)
            return d_o.read(np.uint32, n), d_c.read(np.uint32, 3 * n).reshape(n, 3)
        self.dispatch(d_r.ptr, n, B.RTQ_TRACE_ANY, d_o.ptr, None)
        return d_o.read(np.uint32, n)

    def shadow(self, rays: np.ndarray, hits: np.ndarray, nid: np.ndarray, sun_dir, flags: int) -> np.ndarray:
        n = rays.size
        d_r = self.dev.upload(rays)
        d_h = self.dev.upload(hits)
        d_n = self.dev.upload(np.ascontiguousarray(nid, np.uint32))
        d_m = self.dev.alloc(4 * max(n, 1))
        self.shadow_masks(d_r.ptr, d_h.ptr, d_n.ptr, n, sun_dir, flags, d_m.ptr)
        return d_m.read(np.uint32, n)

    def cull_lights_async(self, frustum: B.Frustum, width: int, height: int, d_offsets: int, d_lists: int,
                          capacity: int, stream=None):
        """rtq_cull_lights without h_total: asynchronous on `stream`, the total lands in
        d_offsets[num_tiles]."""
        B.check_device("rtq_cull_lights",
                       self.dev.lib.rtq_cull_lights(self.handle, C.byref(frustum), width, height, d_offsets, d_lists,
                                                    capacity, None, stream))

    def cull_lights(self, frustum: B.Frustum, width: int, height: int, capacity: int | None = None):
        ntiles = ((width + 15) // 16) * ((height + 15) // 16)
        d_off = self.dev.alloc(4 * (ntiles + 1))
        total = C.c_uint32()
        if capacity is None:
            # first pass with no list storage gives the total
            st = self.dev.lib.rtq_cull_lights(self.handle, C.byref(frustum), width, height, d_off.ptr, None, 0,
                                              C.byref(total), None)
            if st not in (B.RTQ_OK, 4):
                B.check_device("rtq_cull_lights", st)
            capacity = total.value
        d_l = self.dev.alloc(4 * max(1, capacity))
        B.check_device("rtq_cull_lights",
                       self.dev.lib.rtq_cull_lights(self.handle, C.byref(frustum), width, height, d_off.ptr, d_l.ptr,
                                                    capacity, C.byref(total), None))
        offs = d_off.read(np.uint32, ntiles + 1)
        lists = d_l.read(np.uint32, max(1, total.value))[:total.value]
        return offs, lists
