"""Python view of the host builder and workloads (librtq_build.so), used by tests and bench.py.

Method names follow the reference classes they stand for: BVHGeometry::addTriangleList,
BVHBuilder::addTriangleList / addSphereLight / setParameters / build / getBvhGpuSize /
fillGpuBuffer (Renderer/BVHBuilder.h:63-101, Renderer/BVHGeometry.h:17-22).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import bindings as B

# fov of the reference camera pass: 70 * 3.14 / 180 in fp32 (Renderer/raytracingPass.cpp:43)
REFERENCE_FOV = float(np.float32(70.0) * np.float32(3.14) / np.float32(180.0))
TMAX = 100.0


class BVHBuilder:
    def __init__(self, lib=None):
        """lib: another build of librtq_build.so bound by bindings.bind_build_lib (tests); default the
        package's own."""
        self._lib = lib if lib is not None else B.build_lib()
        h = C.c_void_p()
        B.check_build("rtq_builder_create", self._lib.rtq_builder_create(C.byref(h)))
        self._h = h
        self.built = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.rtq_builder_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    # BVHGeometry::addTriangleList
    def add_vertices(self, xyz) -> int:
        a = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        off = C.c_uint32()
        B.check_build("rtq_builder_add_vertices",
                      self._lib.rtq_builder_add_vertices(self._h, B.ptr(a), a.shape[0], C.byref(off)))
        self.built = False
        return off.value

    # BVHBuilder::addTriangleList
    def add_triangle_list(self, vertex_offset: int, indices, color=(0.7, 0.7, 0.7)):
        idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        assert idx.size % 3 == 0
        col = np.asarray(color, dtype=np.float32)
        B.check_build("rtq_builder_add_triangle_list",
                      self._lib.rtq_builder_add_triangle_list(self._h, vertex_offset, idx.size // 3,
                                                              B.ptr(idx), B.ptr(col)))
        self.built = False

    def add_sphere_light(self, pos, radius, color=(1, 1, 1), sphere_radius=0.1):
        p = np.asarray(pos, dtype=np.float32)
        c = np.asarray(color, dtype=np.float32)
        B.check_build("rtq_builder_add_sphere_light",
                      self._lib.rtq_builder_add_sphere_light(self._h, B.ptr(p), radius, B.ptr(c), sphere_radius))
        self.built = False

    def set_parameters(self, max_depth=25, min_obj_per_node=4, min_obj_gain=4,
                       expand_node_volume_threshold=0.25, expand_node_dimension_factor=0.5):
        bp = B.BuildParams(max_depth, min_obj_per_node, min_obj_gain,
                           expand_node_volume_threshold, expand_node_dimension_factor)
        B.check_build("rtq_builder_set_parameters", self._lib.rtq_builder_set_parameters(self._h, C.byref(bp)))
        self.built = False

    # BVHBuilder::addBlas: a bottom-level BVH of one triangle list over this builder's vertices
    def add_blas(self, vertex_offset: int, indices, color=(0.7, 0.7, 0.7)) -> int:
        idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        assert idx.size % 3 == 0
        col = np.asarray(color, dtype=np.float32)
        k = C.c_uint32()
        B.check_build("rtq_builder_add_blas",
                      self._lib.rtq_builder_add_blas(self._h, vertex_offset, idx.size // 3, B.ptr(idx), B.ptr(col),
                                                     C.byref(k)))
        self.built = False
        return k.value

    def set_blas_parameters(self, max_depth=25, min_obj_per_node=4, min_obj_gain=4,
                            expand_node_volume_threshold=0.25, expand_node_dimension_factor=0.5):
        bp = B.BuildParams(max_depth, min_obj_per_node, min_obj_gain,
                           expand_node_volume_threshold, expand_node_dimension_factor)
        B.check_build("rtq_builder_set_blas_parameters",
                      self._lib.rtq_builder_set_blas_parameters(self._h, C.byref(bp)))
        self.built = False

    def generate(self, kind: int, seed: int, param: int = 0):
        B.check_build("rtq_scene_generate", self._lib.rtq_scene_generate(self._h, kind, seed, param))
        self.built = False

    def generate_lights(self, seed: int, count: int, lo=-100.0, hi=100.0, rmin=0.5, rmax=20.0):
        B.check_build("rtq_lights_generate",
                      self._lib.rtq_lights_generate(self._h, seed, count, lo, hi, rmin, rmax))
        self.built = False

    def build(self, threads: int = 1):
        B.check_build("rtq_builder_build", self._lib.rtq_builder_build(self._h, threads))
        self.built = True
        return self.stats()

    def stats(self) -> B.BuildStats:
        s = B.BuildStats()
        B.check_build("rtq_builder_get_stats", self._lib.rtq_builder_get_stats(self._h, C.byref(s)))
        return s

    def gpu_size(self) -> int:
        n = C.c_size_t()
        B.check_build("rtq_builder_gpu_size", self._lib.rtq_builder_gpu_size(self._h, C.byref(n)))
        return n.value

    def fill_gpu_buffer(self):
        """-> (blob as uint8 array, Ranges)"""
        blob = np.zeros(self.gpu_size(), dtype=np.uint8)
        r = B.Ranges()
        B.check_build("rtq_builder_fill_gpu_buffer",
                      self._lib.rtq_builder_fill_gpu_buffer(self._h, B.ptr(blob), blob.size, C.byref(r)))
        return blob, r

    def positions(self) -> np.ndarray:
        p = C.c_void_p()
        n = C.c_uint32()
        B.check_build("rtq_builder_positions", self._lib.rtq_builder_positions(self._h, C.byref(p), C.byref(n)))
        if n.value == 0:
            return np.zeros((0, 3), np.float32)
        buf = (C.c_float * (3 * n.value)).from_address(p.value)
        return np.frombuffer(buf, dtype=np.float32).reshape(-1, 3).copy()

    def input_hash(self) -> int:
        h = C.c_uint64()
        B.check_build("rtq_builder_input_hash", self._lib.rtq_builder_input_hash(self._h, C.byref(h)))
        return h.value

    def save_cache(self, path: str):
        B.check_build("rtq_cache_save", self._lib.rtq_cache_save(self._h, os.fsencode(path)))

    def scene_box(self):
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        B.check_build("rtq_builder_scene_box", self._lib.rtq_builder_scene_box(self._h, B.ptr(mn), B.ptr(mx)))
        return mn, mx

    def surface_rays(self, seed: int, count: int, tmax: float = TMAX) -> np.ndarray:
        out = np.zeros(count, dtype=B.RAY_DTYPE)
        B.check_build("rtq_surface_rays", self._lib.rtq_surface_rays(self._h, seed, count, tmax, B.ptr(out)))
        return out

    def bounce_rays(self, seed: int, rays: np.ndarray, hits: np.ndarray, tmax: float = TMAX) -> np.ndarray:
        assert rays.dtype == B.RAY_DTYPE and hits.dtype == B.HIT_DTYPE and rays.size == hits.size
        out = np.zeros(rays.size, dtype=B.RAY_DTYPE)
        B.check_build("rtq_bounce_rays", self._lib.rtq_bounce_rays(self._h, seed, B.ptr(rays), B.ptr(hits),
                                                                   rays.size, tmax, B.ptr(out)))
        return out


def camera_frustum(pos, direction, up=(0, 0, 1), fov_y=REFERENCE_FOV, aspect=1.0) -> B.Frustum:
    lib = B.build_lib()
    p = np.asarray(pos, np.float32)
    d = np.asarray(direction, np.float32)
    u = np.asarray(up, np.float32)
    f = B.Frustum()
    B.check_build("rtq_camera_frustum", lib.rtq_camera_frustum(B.ptr(p), B.ptr(d), B.ptr(u), fov_y, aspect,
                                                               C.byref(f)))
    return f


def frustum_rays(f: B.Frustum, width: int, height: int, seed: int, tmax: float = TMAX) -> np.ndarray:
    lib = B.build_lib()
    out = np.zeros(width * height, dtype=B.RAY_DTYPE)
    B.check_build("rtq_frustum_rays", lib.rtq_frustum_rays(C.byref(f), width, height, seed, tmax, B.ptr(out)))
    return out


def load_cache(path: str, expected_hash: int = 0):
    """(blob uint8 array, Ranges, positions (n, 3) float32, key) from a scene cache file, or None when
    the file is missing or was built from other inputs. A corrupt file raises."""
    lib = B.build_lib()
    h = C.c_void_p()
    st = lib.rtq_cache_load(os.fsencode(path), expected_hash, C.byref(h))
    if st == B.RTQ_ERR_CACHE_MISS:
        return None
    B.check_build("rtq_cache_load", st)
    try:
        blob_p, nbytes, r = C.c_void_p(), C.c_size_t(), B.Ranges()
        pos_p, nv, key = C.c_void_p(), C.c_uint32(), C.c_uint64()
        B.check_build("rtq_cached_scene_view", lib.rtq_cached_scene_view(
            h, C.byref(blob_p), C.byref(nbytes), C.byref(r), C.byref(pos_p), C.byref(nv), C.byref(key)))
        blob = np.ctypeslib.as_array((C.c_uint8 * nbytes.value).from_address(blob_p.value)).copy() \
            if nbytes.value else np.zeros(0, np.uint8)
        pos = np.ctypeslib.as_array((C.c_float * (3 * nv.value)).from_address(pos_p.value)).reshape(-1, 3).copy() \
            if nv.value else np.zeros((0, 3), np.float32)
        return blob, r, pos, key.value
    finally:
        lib.rtq_cached_scene_free(h)


def probe_directions(seed: int, count: int = 64) -> np.ndarray:
    """Stratified unit directions (count, 4) as the probe pass samples them."""
    out = np.zeros((count, 4), np.float32)
    B.check_build("rtq_probe_directions", B.build_lib().rtq_probe_directions(seed, count, B.ptr(out)))
    return out


def workload(config: int, scale: float = 1.0, threads: int = 8):
    """Build one of the benchmark scenes. scale < 1 shrinks the triangle/light count for tests.

    config 1: closed box, 16 384 tris; 256x256 primary rays from (5,1,5) looking +y.
    config 2: 4096 UV spheres x 256 = 1 048 576 tris; 1920x1080 primary rays from the origin
              looking +x; one point light at (0,0,150) with radius 1e4.
    config 3: 3 Mi-triangle Gaussian-mixture soup with 4096 point lights.
    Returns (builder, frustum, width, height).
    """
    b = BVHBuilder()
    if config == 1:
        b.generate(B.RTQ_SCENE_BOX, 1, 0)
        w = h = 256
        f = camera_frustum((5, 1, 5), (0, 1, 0), (0, 0, 1), REFERENCE_FOV, 1.0)
    elif config == 2:
        b.generate(B.RTQ_SCENE_SPHERES, 2, max(1, int(round(4096 * scale))))
        b.add_sphere_light((0, 0, 150), 1e4, (1, 1, 1), 0.1)
        w, h = 1920, 1080
        f = camera_frustum((0, 0, 0), (1, 0, 0), (0, 0, 1), REFERENCE_FOV, w / h)
    elif config == 3:
        b.generate(B.RTQ_SCENE_SOUP, 3, max(1, int(round(3 * 1024 * 1024 * scale))))
        b.generate_lights(3, max(1, int(round(4096 * min(1.0, scale * 4)))), -100, 100, 0.5, 20)
        w, h = 1920, 1080
        f = camera_frustum((0, 0, 0), (1, 0, 0), (0, 0, 1), REFERENCE_FOV, w / h)
    else:
        raise ValueError(config)
    b.build(threads)
    return b, f, w, h


# ---- many-scene workloads (the batched and served forms) ------------------------------------------

def loguniform(rng, lo: float, hi: float, n: int) -> np.ndarray:
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def sphere_builders(sizes, seed0: int):
    """One unbuilt builder per triangle count: a soup of ceil(tris / 256) seeded UV spheres of 256
    triangles each (the spheres generator), seeds seed0 + i."""
    out = []
    for i, tris in enumerate(sizes):
        b = BVHBuilder()
        b.generate(B.RTQ_SCENE_SPHERES, seed0 + i, max(1, int(math.ceil(tris / 256))))
        out.append(b)
    return out


def build_many(builders, threads: int = 16, cache_dir: str | None = None, progress=None):
    """Builds every builder, up to `threads` at once and largest first. A scene gets one host thread,
    or a few if it holds a large share of all triangles (so the largest scenes do not finish alone at
    the end); results do not depend on the thread count. With cache_dir, a scene whose inputs match a
    cached file is loaded from it instead and built scenes are stored there. Returns one
    (blob, Ranges, positions) per builder, in order."""
    from concurrent.futures import ThreadPoolExecutor

    tris = [builders[i].stats().num_triangles for i in range(len(builders))]
    total = max(1, sum(tris))
    order = sorted(range(len(builders)), key=lambda i: -tris[i])
    out = [None] * len(builders)
    done = [0]

    def one(i):
        b = builders[i]
        path = None
        if cache_dir:
            os.makedirs(cache_dir, exist_ok=True)
            key = b.input_hash()
            path = os.path.join(cache_dir, f"{key:016x}.rtqbvh")
            hit = load_cache(path, key)
            if hit is not None:
                out[i] = hit[:3]
                return
        b.build(max(1, min(threads, int(4 * threads * tris[i] / total))))
        if path:
            b.save_cache(path)
        blob, ranges = b.fill_gpu_buffer()
        out[i] = (blob, ranges, b.positions())
        done[0] += 1
        if progress and done[0] % 32 == 0:
            progress(f"built {done[0]} of {len(builders)} scenes")

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(one, order))
    return out


def primary_rays(n: int, seed: int) -> np.ndarray:
    """n jittered primary rays from the origin looking +x (square frame, reference field of view)."""
    side = int(np.ceil(np.sqrt(n)))
    f = camera_frustum((0, 0, 0), (1, 0, 0), (0, 0, 1), REFERENCE_FOV, 1.0)
    return np.ascontiguousarray(frustum_rays(f, side, side, seed=seed)[:n])


def batch_workload(scale: float = 1.0, slots: int = 256):
    """The batched form's workload: `slots` scenes of LogU[1K, 2M * scale] triangles (seeds 1000 + i)
    and per item LogU[4K, 1M] rays, primary rays for even items and diffuse surface rays for odd ones.
    Returns (unbuilt builders, ray counts); item_rays() makes an item's rays."""
    rng = np.random.default_rng(4)
    sizes = loguniform(rng, 1000, 2e6 * scale, slots)
    counts = loguniform(rng, 4096, 1 << 20, slots).astype(np.int64)
    return sphere_builders(sizes, 1000), counts


def batch_item_rays(builder, item: int, n: int) -> np.ndarray:
    return primary_rays(n, 10 + item) if item % 2 == 0 else builder.surface_rays(seed=10 + item, count=n)


def serve_workload(scale: float = 1.0, slots: int = 64, requests: int = 1000):
    """The served form's workload: `slots` resident scenes of LogU[10K, 3M * scale] triangles (seeds
    2000 + i) and a request stream of LogU[1K, 2M] rays each, half closest and half any hit, on uniformly
    drawn scenes. Returns (unbuilt builders, rng, request sizes, request scenes, any-hit flags); the rng
    continues for the arrival times."""
    rng = np.random.default_rng(5)
    sizes = loguniform(rng, 1e4, 3e6 * scale, slots)
    builders = sphere_builders(sizes, 2000)
    req_sizes = loguniform(rng, 1000, 2e6, requests).astype(np.int64)
    req_scene = rng.integers(0, slots, requests)
    req_any = rng.random(requests) < 0.5
    return builders, rng, req_sizes, req_scene, req_any


def two_level_scene(seed: int, n_blas: int = 24, tris_per_blas: int = 300, own_triangles: int = 40,
                    spread: float = 30.0):
    """A two-level scene: `n_blas` BLAS instances, each a cluster of small random triangles around its
    own centre in [-spread, spread]^3, and `own_triangles` large triangles of the top level on the plane
    z = -spread - 10 + 0.1 x (the reference's top-level traversal never tests those; the plane is tilted
    because a box flat in z at |z| >= 16 stays flat after the builder's 1e-6 widening, and CollideBox
    never accepts a flat box). The top level uses the
    application's TLAS parameters (main.cpp:203-206: 6, 6, 1.0 and the default depth and factor), the
    BLASes its BLAS parameters (main.cpp:197-201, the builder's defaults). Unbuilt."""
    rng = np.random.default_rng(seed)
    b = BVHBuilder()
    b.set_parameters(25, 6, 6, 1.0, 0.225)
    centres = rng.uniform(-spread, spread, (n_blas, 3)).astype(np.float32)
    for k in range(n_blas):
        c = centres[k] + rng.normal(size=(tris_per_blas, 1, 3)) * rng.uniform(0.5, 3.0)
        v = (c + rng.normal(size=(tris_per_blas, 3, 3)) * 0.4).astype(np.float32).reshape(-1, 3)
        off = b.add_vertices(v)
        b.add_blas(off, np.arange(3 * tris_per_blas), color=rng.uniform(0.1, 0.9, 3))
    if own_triangles:
        z = -spread - 10.0
        c = rng.uniform(-spread, spread, (own_triangles, 1, 2))
        xy = c + rng.normal(size=(own_triangles, 3, 2)) * 4.0
        v = np.concatenate([xy, z + 0.1 * xy[..., :1]], axis=2).astype(np.float32).reshape(-1, 3)
        off = b.add_vertices(v)
        b.add_triangle_list(off, np.arange(3 * own_triangles))
    return b


def two_level_rays(builder, seed: int, n: int) -> np.ndarray:
    """(n, 8) float32 rays (rtq_ray words, t limit 1e4) for a built two_level_scene, four kinds in equal
    parts: from afar toward the BLAS clusters, from inside BLAS header boxes toward other clusters, down
    onto the top level's own triangles, and random ones."""
    blob, r = builder.fill_gpu_buffer()
    hf = blob[r.blas.offset:r.blas.offset + r.blas.size].view(np.float32).reshape(-1, 8)
    pos = builder.positions()
    rng = np.random.default_rng(seed)
    q = n // 4
    rays = np.zeros((n, 8), np.float32)
    rays[:, 3] = 1e4
    lo, hi = pos.min(0), pos.max(0)
    span = float((hi - lo).max())
    centres = (hf[:, 0:3] + hf[:, 4:7]) / 2
    org = rng.normal(size=(q, 3)) * span
    tgt = centres[rng.integers(0, len(hf), q)] + rng.normal(size=(q, 3))
    rays[:q, 0:3], rays[:q, 4:7] = org, tgt - org
    k = rng.integers(0, len(hf), q)
    org = hf[k, 0:3] + rng.uniform(0.05, 0.95, (q, 3)) * (hf[k, 4:7] - hf[k, 0:3])
    tgt = centres[rng.integers(0, len(hf), q)] + rng.normal(size=(q, 3))
    rays[q:2 * q, 0:3], rays[q:2 * q, 4:7] = org, tgt - org + rng.normal(size=(q, 3)) * 0.1
    org = np.column_stack([rng.uniform(lo[0], hi[0], q), rng.uniform(lo[1], hi[1], q), np.full(q, lo[2] + 8.0)])
    rays[2 * q:3 * q, 0:3], rays[2 * q:3 * q, 4:7] = org, np.column_stack([rng.normal(size=(q, 2)) * 0.2, -np.ones(q)])
    m = n - 3 * q
    rays[3 * q:, 0:3], rays[3 * q:, 4:7] = rng.normal(size=(m, 3)) * (span / 1.5), rng.normal(size=(m, 3))
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7], axis=1, keepdims=True)
    return np.ascontiguousarray(rays)
