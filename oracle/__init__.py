"""Scalar CPU oracle for the ray-query path: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this package; the
product path (tim_raytracing_amd) never does. The C restatement lives in rtq_oracle.c (built by
the Makefile beside it into liborc.so); this module is its ctypes wrapper.

Parity pins (what checks each oracle function, see tests/test_oracle_pins.py):
  orc_collide_triangle  closed forms on exactly representable triangles, FP64 bound, planted-mistake
  orc_collide_box       closed forms, zero-direction/NaN cases
  traverse (closest)    bitstack == stack form; single-leaf scenes == brute force over leaf order;
                        multi-leaf scenes == brute-force minimum within the 1e-5 shrink window;
                        invariance under node renumbering; planted tie-break/shrink mistakes fail
  traverse (any)        == (closest raw t < tmax) on every ray; order independence
  shadow masks          scenes with a known occluder / no occluder
  probe masks           closed forms: a floor and two walls give, per interval of the hit's x, which
                        corner segments are blocked; planted mistakes 4 and 5 fail
  probe field tracing   tiled probe numbering inverted by the reference's getLightProbIndex; axis
                        rays from probes in a closed box hit the wall at the known distance
                        (planted mistake 6 fails)
  light culling         closed forms (light on the view axis, light behind the camera, huge radius);
                        lights inside and outside the quarter-step widening band of a tile
                        (planted mistake 7 fails)
  camera rays           closed forms of the pinhole model; the reference storage row
                        16*ceil(H/16) - y on a height that is not a multiple of 16 (planted mistake 8
                        fails)
  FP64 hits             closed forms on axis-aligned dyadic triangles
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RTQ_ORACLE_LIB names another build of the oracle (the tests' sanitizer build)
LIB = os.environ.get("RTQ_ORACLE_LIB") or os.path.join(HERE, "liborc.so")
REF_TRIBOX = os.path.join(HERE, "_ref", "libtribox_ref.so")
PORT_TRIBOX = os.path.join(HERE, "libtribox_port.so")

P = C.c_void_p
U32 = C.c_uint32

_SIGS = {
    "orc_collide_box": ([P, P, P, P, C.c_float], C.c_float),
    "orc_collide_triangle": ([P, P, P, P, P, C.c_float, P, P], C.c_float),
    "orc_intersect_f64": ([P, P, P, P, P, P, P, P], C.c_int),
    "orc_hits_f64": ([P, C.c_size_t, P, P, U32, P, U32, P, P, P, P], C.c_int),
    "orc_trace_closest": ([P, C.c_size_t, P, P, U32, P, U32, P, P, P, P, P, P, P, C.c_int, C.c_int], C.c_int),
    "orc_trace_any": ([P, C.c_size_t, P, P, U32, P, U32, P, P, C.c_int, C.c_int], C.c_int),
    "orc_trace_any_tlas": ([P, C.c_size_t, P, P, U32, P, U32, P, C.c_int], C.c_int),
    "orc_brute_closest": ([P, C.c_size_t, P, P, U32, P, U32, C.c_float, P, P, P], C.c_int),
    "orc_shadow_masks": ([P, C.c_size_t, P, P, U32, P, P, P, U32, P, U32, P, C.c_int], C.c_int),
    "orc_tile_planes": ([P, U32, U32, U32, U32, P], None),
    "orc_cull_lights": ([P, U32, P, U32, U32, P, P, C.c_uint64], C.c_uint64),
    "orc_camera_rays": ([P, U32, U32, C.c_int, P], None),
    "orc_trace_record": ([P, C.c_size_t, P, P, U32, P, U32, U32, C.c_int, P, U32, P, P, P, C.c_int], C.c_int),
    "orc_trace_probes": ([P, C.c_size_t, P, P, U32, P, P, P, P, U32, P, C.c_int], C.c_int),
    "orc_probe_position": ([P, P, U32, P], None),
}

_cache = {}


def build():
    """Compile the oracle (and the reference SAT test where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(plant: int = 0):
    name = LIB if plant == 0 else os.path.join(HERE, f"liborc_plant{plant}.so")
    if name not in _cache:
        if not os.path.exists(name):
            build()
        L = C.CDLL(name)
        for fn, (a, r) in _SIGS.items():
            f = getattr(L, fn)
            f.argtypes = a
            f.restype = r
        _cache[name] = L
    return _cache[name]


def _p(a):
    return None if a is None else a.ctypes.data


def _ranges_array(ranges) -> np.ndarray:
    names = ("triangles", "primitives", "materials", "lights", "nodes", "leaf_data", "blas")
    out = np.zeros(14, np.uint32)
    for i, n in enumerate(names):
        r = getattr(ranges, n)
        out[2 * i] = r.offset
        out[2 * i + 1] = r.size
    return out


class Scene:
    """Blob + positions as the shaders see them."""

    def __init__(self, blob: np.ndarray, ranges, positions: np.ndarray):
        self.blob = np.ascontiguousarray(blob, dtype=np.uint8)
        self.ranges = _ranges_array(ranges) if not isinstance(ranges, np.ndarray) else ranges.astype(np.uint32)
        self.pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        self.nv = self.pos.size // 3

    def lights(self) -> np.ndarray:
        off, size = int(self.ranges[6]), int(self.ranges[7])
        n = size // 56
        return np.frombuffer(self.blob[off:off + 56 * n].tobytes(), dtype=np.float32).reshape(n, 14).copy()

    def trace_closest(self, rays: np.ndarray, form: int = 0, threads: int = 8, plant: int = 0):
        n = rays.size
        rays = np.ascontiguousarray(rays)
        out = {
            "tri_id": np.zeros(n, np.uint32), "t": np.zeros(n, np.float32), "u": np.zeros(n, np.float32),
            "v": np.zeros(n, np.float32), "nid": np.zeros(n, np.uint32), "slot": np.zeros(n, np.uint32),
            "counters": np.zeros((n, 3), np.uint32),
        }
        rc = lib(plant).orc_trace_closest(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv,
                                          _p(rays), n, _p(out["tri_id"]), _p(out["t"]), _p(out["u"]),
                                          _p(out["v"]), _p(out["nid"]), _p(out["slot"]), _p(out["counters"]),
                                          form, threads)
        assert rc == 0, "oracle rejected the blob"
        return out

    def trace_any(self, rays: np.ndarray, form: int = 0, threads: int = 8, plant: int = 0):
        n = rays.size
        rays = np.ascontiguousarray(rays)
        occ = np.zeros(n, np.uint32)
        cnt = np.zeros((n, 3), np.uint32)
        rc = lib(plant).orc_trace_any(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv,
                                      _p(rays), n, _p(occ), _p(cnt), form, threads)
        assert rc == 0
        return occ, cnt

    def trace_any_tlas(self, rays: np.ndarray, threads: int = 8, plant: int = 0) -> np.ndarray:
        """Two-level any hit (traverseTlasFast with tlas_collide_fast): 1 = occluded. rays: rtq_ray
        records or an (n, 8) float32 array of the same words."""
        if rays.dtype.names is None:
            assert rays.dtype == np.float32 and rays.ndim == 2 and rays.shape[1] == 8
        n = rays.size if rays.dtype.names else rays.shape[0]
        occ = np.zeros(n, np.uint32)
        rc = lib(plant).orc_trace_any_tlas(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv,
                                           _p(np.ascontiguousarray(rays)), n, _p(occ), threads)
        assert rc == 0, "oracle rejected the two-level blob"
        return occ

    def brute_closest(self, rays: np.ndarray, window: float = 1e-5):
        n = rays.size
        rays = np.ascontiguousarray(rays)
        tri = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        near = np.zeros(n, np.uint32)
        rc = lib().orc_brute_closest(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv,
                                     _p(rays), n, window, _p(tri), _p(t), _p(near))
        assert rc == 0
        return tri, t, near

    def hits_f64(self, rays: np.ndarray, tri_id: np.ndarray):
        """FP64 (t, u, v) of each ray against its recorded hit triangle; NaN for misses."""
        n = rays.size
        out = [np.zeros(n, np.float64) for _ in range(3)]
        rc = lib().orc_hits_f64(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv,
                                _p(np.ascontiguousarray(rays)), n, _p(np.ascontiguousarray(tri_id, np.uint32)),
                                *[_p(o) for o in out])
        assert rc == 0, "bad triangle id or blob"
        return tuple(out)

    def shadow_masks(self, rays, hit_t, hit_nid, sun_dir, flags, threads: int = 8):
        n = rays.size
        masks = np.zeros(n, np.uint32)
        sd = np.asarray(sun_dir, np.float32)
        rc = lib().orc_shadow_masks(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv,
                                    _p(np.ascontiguousarray(rays)), _p(np.ascontiguousarray(hit_t, np.float32)),
                                    _p(np.ascontiguousarray(hit_nid, np.uint32)), n, _p(sd), flags, _p(masks),
                                    threads)
        assert rc == 0
        return masks

    def trace_record(self, frustum, width: int, height: int, reference_rows: bool = False, sun_dir=None,
                     shadow_flags: int = 0, lpf=None, threads: int = 8, plant: int = 0) -> np.ndarray:
        """The camera pass's TRACING_STEP record, (H*W, 8) uint32; lpf = (min3, max3, res3) or None."""
        fr = frustum_array(frustum)
        rec = np.zeros((width * height, 8), np.uint32)
        sd = np.asarray(sun_dir if sun_dir is not None else (0, 0, 0), np.float32)
        box = res = None
        if lpf is not None:
            box = np.array(list(lpf[0]) + list(lpf[1]), np.float32)
            res = np.array(lpf[2], np.uint32)
        rc = lib(plant).orc_trace_record(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv, _p(fr),
                                    width, height, int(reference_rows), _p(sd), shadow_flags, _p(box), _p(res),
                                    _p(rec), threads)
        assert rc == 0
        return rec

    def trace_probes(self, lpf, dirs, sun_dir=None, shadow_flags: int = 0, threads: int = 8,
                     plant: int = 0) -> np.ndarray:
        """The light-probe field's TRACING_STEP records, (probes*64, 8) uint32, sample = probe*64 + ray."""
        box = np.array(list(lpf[0]) + list(lpf[1]), np.float32)
        res = np.array(lpf[2], np.uint32)
        n = int(res.prod())
        rec = np.zeros((n * 64, 8), np.uint32)
        d = np.ascontiguousarray(dirs, np.float32).reshape(64, 4)
        sd = np.asarray(sun_dir if sun_dir is not None else (0, 0, 0), np.float32)
        rc = lib(plant).orc_trace_probes(_p(self.blob), self.blob.size, _p(self.ranges), _p(self.pos), self.nv, _p(box),
                                    _p(res), _p(d), _p(sd), shadow_flags, _p(rec), threads)
        assert rc == 0
        return rec

    def cull_lights(self, frustum, width: int, height: int):
        lights = self.lights()
        return cull_lights(lights, frustum, width, height)


def frustum_array(f) -> np.ndarray:
    if isinstance(f, np.ndarray):
        return f.astype(np.float32)
    return np.array(list(f.camera_pos) + list(f.corner00) + list(f.corner10) + list(f.corner01), np.float32)


def cull_lights(lights: np.ndarray, frustum, width: int, height: int, plant: int = 0):
    """lights: (n, 14) float32 PackedLight records. -> (offsets, lists)"""
    fr = frustum_array(frustum)
    ntiles = ((width + 15) // 16) * ((height + 15) // 16)
    offs = np.zeros(ntiles + 1, np.uint32)
    L = np.ascontiguousarray(lights, np.float32)
    total = lib(plant).orc_cull_lights(_p(L), L.shape[0], _p(fr), width, height, _p(offs), None, 0)
    lists = np.zeros(max(1, total), np.uint32)
    lib(plant).orc_cull_lights(_p(L), L.shape[0], _p(fr), width, height, _p(offs), _p(lists), total)
    return offs, lists[:total]


def camera_rays(frustum, width: int, height: int, reference_rows: bool = False, plant: int = 0) -> np.ndarray:
    """Primary rays exactly as the camera pass generates them (rtq_ray records, y*W + x)."""
    fr = frustum_array(frustum)
    out = np.zeros(width * height * 8, np.float32)
    lib(plant).orc_camera_rays(_p(fr), width, height, int(reference_rows), _p(out))
    return out.view(np.dtype([("org", "<f4", 3), ("tmax", "<f4"), ("dir", "<f4", 3), ("pad", "<u4")]))


def probe_position(lpf, index: int) -> np.ndarray:
    box = np.array(list(lpf[0]) + list(lpf[1]), np.float32)
    res = np.array(lpf[2], np.uint32)
    out = np.zeros(3, np.float32)
    lib().orc_probe_position(_p(box), _p(res), index, _p(out))
    return out


def tile_planes(frustum, width, height, tx, ty) -> np.ndarray:
    fr = frustum_array(frustum)
    out = np.zeros(16, np.float32)
    lib().orc_tile_planes(_p(fr), width, height, tx, ty, _p(out))
    return out.reshape(4, 4)


def collide_box(o, d, mn, mx, tmax):
    a = [np.asarray(x, np.float32) for x in (o, d, mn, mx)]
    return float(lib().orc_collide_box(*[_p(x) for x in a], tmax))


def collide_triangle(o, d, p0, p1, p2, tmax, plant: int = 0):
    a = [np.asarray(x, np.float32) for x in (o, d, p0, p1, p2)]
    u = np.zeros(1, np.float32)
    v = np.zeros(1, np.float32)
    t = lib(plant).orc_collide_triangle(*[_p(x) for x in a], tmax, _p(u), _p(v))
    return float(t), float(u[0]), float(v[0])


def intersect_f64(o, d, p0, p1, p2):
    a = [np.asarray(x, np.float32) for x in (o, d, p0, p1, p2)]
    t = np.zeros(1, np.float64)
    u = np.zeros(1, np.float64)
    v = np.zeros(1, np.float64)
    ok = lib().orc_intersect_f64(*[_p(x) for x in a], _p(t), _p(u), _p(v))
    return bool(ok), float(t[0]), float(u[0]), float(v[0])
