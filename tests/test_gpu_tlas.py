"""GPU parity of two-level any hit: rtq_dispatch_rays(RTQ_TRACE_ANY) on a two-level scene (the any-hit
kernels entering BLAS instances at top-level leaves) against the oracle's traverseTlasFast /
tlas_collide_fast restatement on the same blob and rays, in every dispatch mode.

Bar: identical occlusion flags on every ray. Scenes: 24 BLAS clusters with top-level triangles, a top
level that is one leaf, BLASes that are one leaf each, and 1024 BLAS instances of 1000 triangles (about
one million triangles). The other queries reject two-level scenes.
"""
import numpy as np
import pytest

import oracle
from tests import gpu_util as G
from tests import test_tlas as T
from tim_raytracing_amd import bindings as B
from tim_raytracing_amd import device as D
from tim_raytracing_amd import host as H

pytestmark = pytest.mark.gpu


def _print(msg):
    print(msg, flush=True)


def _records(rays):
    return np.ascontiguousarray(rays).view(B.RAY_DTYPE).reshape(-1)


def _scene(seed, **kw):
    b = H.two_level_scene(seed, **kw)
    b.build(16)
    blob, r = b.fill_gpu_buffer()
    pos = b.positions()
    return b, D.Scene(G.dev(), blob, r, pos), oracle.Scene(blob, r, pos)


@pytest.mark.parametrize("seed,kw", [(3, {}), (5, dict(n_blas=4, tris_per_blas=150, own_triangles=2)),
                                     (7, dict(n_blas=40, tris_per_blas=3))])
def test_two_level_any_hit_matches_oracle(seed, kw):
    _print(f"case: two-level any hit, seed {seed}, {kw or 'default scene'}, 40000 rays")
    b, g, o = _scene(seed, **kw)
    rays = _records(T.scene_rays(b, seed, 40000))
    got = g.trace_any(rays)
    want = o.trace_any_tlas(rays, threads=16)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} rays differ, first {bad[:5]}"
    assert want.sum() > 100 and want.mean() < 0.95  # both answers well represented


@pytest.mark.parametrize("persistent,sort", [(0, 1), (1, 0), (1, 1), (2, 0), (2, 1)])
def test_two_level_dispatch_modes_identical(persistent, sort):
    _print(f"case: two-level any hit, dispatch persistent={persistent} sort={sort}, 24 BLAS, 40000 rays")
    d = G.dev()
    keep = (d.get_option(B.RTQ_OPTION_PERSISTENT), d.get_option(B.RTQ_OPTION_SORT_RAYS))
    d.set_option(B.RTQ_OPTION_PERSISTENT, persistent)
    d.set_option(B.RTQ_OPTION_SORT_RAYS, sort)
    try:
        b, g, o = _scene(3)
        rays = _records(T.scene_rays(b, 9, 40000))
        want = o.trace_any_tlas(rays, threads=16)
        for n in (rays.size, 37, 1):
            got = g.trace_any(np.ascontiguousarray(rays[:n]))
            bad = np.flatnonzero(got != want[:n])
            assert bad.size == 0, f"{n} rays: {bad.size} differ, first {bad[:5]}"
    finally:
        d.set_option(B.RTQ_OPTION_PERSISTENT, keep[0])
        d.set_option(B.RTQ_OPTION_SORT_RAYS, keep[1])


def test_two_level_any_hit_million_triangles():
    _print("case: two-level any hit, 1024 BLAS x 1000 triangles, 1 Mi rays")
    b, g, o = _scene(11, n_blas=1024, tris_per_blas=1000, own_triangles=200, spread=100.0)
    rays = _records(T.scene_rays(b, 12, 1 << 20))
    got = g.trace_any(rays)
    want = o.trace_any_tlas(rays, threads=16)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} rays differ, first {bad[:5]}"


def test_two_level_scene_rejects_other_queries():
    _print("case: two-level scene, closest hit / counted / camera / batch / server rejected")
    b, g, _ = _scene(3)
    rays = _records(T.scene_rays(b, 1, 64))
    with pytest.raises(B.RtqError):
        g.trace_closest(rays)
    with pytest.raises(B.RtqError):
        g.trace_any(rays, counted=True)
    f = H.camera_frustum((0, 0, 60), (0, 0, -1), (0, 1, 0), H.REFERENCE_FOV, 1.0)
    with pytest.raises(B.RtqError):
        g.trace_camera(f, 16, 16)
    d_r = G.dev().upload(rays)
    d_o = G.dev().alloc(4 * rays.size)
    with pytest.raises(B.RtqError):
        D.Batch(G.dev(), [g], [(0, B.RTQ_TRACE_ANY, d_r.ptr, rays.size, d_o.ptr, None)])
    with pytest.raises(B.RtqError):
        D.Server(G.dev(), [g], 1 << 26)
    # the scene still answers any-hit queries after the rejected calls
    assert np.array_equal(g.trace_any(rays), oracle.Scene(*b.fill_gpu_buffer(), b.positions()).trace_any_tlas(rays))
