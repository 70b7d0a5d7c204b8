"""Shared helpers for the GPU tests (scene builds are cached per process)."""
import functools

import numpy as np

import oracle
from tim_raytracing_amd import bindings as B
from tim_raytracing_amd import device as D
from tim_raytracing_amd import host as H


@functools.lru_cache(maxsize=None)
def dev():
    return D.Device(0)


@functools.lru_cache(maxsize=None)
def scene(config: int, scale: float = 1.0):
    b, f, w, h = H.workload(config, scale=scale, threads=16)
    blob, ranges = b.fill_gpu_buffer()
    pos = b.positions()
    return {
        "builder": b, "frustum": f, "w": w, "h": h, "blob": blob, "ranges": ranges, "pos": pos,
        "gpu": D.Scene(dev(), blob, ranges, pos), "orc": oracle.Scene(blob, ranges, pos),
    }


def assert_closest_equal(g, o, label=""):
    """Bit-identical closest-hit records: triangle id, stored t, u, v and the hit leaf's node id."""
    for k in ("tri_id", "nid"):
        bad = np.flatnonzero(g[k] != o[k])
        assert bad.size == 0, f"{label} {k} differs on {bad.size} rays, first {bad[:5]}"
    for k in ("t", "u", "v"):
        gb = np.asarray(g[k], np.float32).view(np.uint32)
        ob = np.asarray(o[k], np.float32).view(np.uint32)
        bad = np.flatnonzero(gb != ob)
        assert bad.size == 0, f"{label} {k} bits differ on {bad.size} rays, first {bad[:5]}"


def tile_order(width: int, height: int, t: int = 8) -> np.ndarray:
    """Pixel indices y*width + x listed in t x t tiles (tiles row-major, pixels row-major inside a
    tile), the ray order of the benchmark's primary buffer."""
    assert width % t == 0 and height % t == 0
    y, x = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    idx = (y * width + x).reshape(height // t, t, width // t, t).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(idx.reshape(-1))


@functools.lru_cache(maxsize=None)
def batch_pool():
    """The batched form's full workload (256 scenes of LogU[1K, 2M] triangles), built once per process
    on 16 host threads: (builders, rays per item, [(blob, ranges, positions)] per scene)."""
    import time

    builders, counts = H.batch_workload()
    t0 = time.time()
    pool = H.build_many(builders, threads=16, progress=lambda m: print(m, flush=True))
    print(f"built {len(pool)} scenes, {sum(b.stats().num_triangles for b in builders)} triangles, in "
          f"{time.time() - t0:.0f} s", flush=True)
    return builders, counts, pool


@functools.lru_cache(maxsize=None)
def serve_pool(requests: int):
    """The served form's full workload (64 resident scenes of LogU[10K, 3M] triangles) and a stream of
    `requests` requests, built once per process on 16 host threads."""
    import time

    builders, rng, sizes, scenes, any_hit = H.serve_workload(requests=requests)
    t0 = time.time()
    pool = H.build_many(builders, threads=16, progress=lambda m: print(m, flush=True))
    print(f"built {len(pool)} scenes, {sum(b.stats().num_triangles for b in builders)} triangles, in "
          f"{time.time() - t0:.0f} s", flush=True)
    return builders, rng, sizes, scenes, any_hit, pool
