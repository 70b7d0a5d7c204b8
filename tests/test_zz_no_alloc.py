"""rtq_cull_lights_noalloc (the culling entry point that allocates nothing) against rtq_cull_lights on
the same inputs, on two calls in a row (the second smaller, so it runs over what the first left in the
shared scratch array), plus truncation at capacity and the size limit. Lists must be identical; the
oracle checks both."""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests import gpu_util as G
from tim_raytracing_amd import bindings as B
from tim_raytracing_amd import device as D
from tim_raytracing_amd import host as H

pytestmark = pytest.mark.gpu

POISON = 0xDEADBEEF


def _scene(num_lights, seed):
    b = H.BVHBuilder()
    b.generate(B.RTQ_SCENE_BOX, 1, 0)
    b.generate_lights(seed, num_lights, -100, 100, 0.5, 20)
    b.build(16)
    blob, r = b.fill_gpu_buffer()
    pos = b.positions()
    return D.Scene(G.dev(), blob, r, pos), oracle.Scene(blob, r, pos)


def _noalloc(scene, f, w, h, capacity):
    dev = G.dev()
    ntiles = ((w + 15) // 16) * ((h + 15) // 16)
    d_off, d_l = dev.alloc(4 * (ntiles + 1)), dev.alloc(4 * max(1, capacity))
    d_off.write(np.full(ntiles + 1, POISON, np.uint32))
    d_l.write(np.full(max(1, capacity), POISON, np.uint32))
    stream = dev.stream()
    B.check_device("rtq_cull_lights_noalloc", dev.lib.rtq_cull_lights_noalloc(
        scene.handle, C.byref(f), w, h, d_off.ptr, d_l.ptr, capacity, stream.ptr))
    stream.synchronize()
    offs = d_off.read(np.uint32, ntiles + 1)
    lists = d_l.read(np.uint32, max(1, capacity))
    return offs, lists


def test_noalloc_culling_equals_allocating_entry_point():
    cases = [(4096, 3, 1920, 1080), (700, 5, 640, 360)]  # the second call is the smaller one
    for num_lights, seed, w, h in cases:
        print(f"case: no-alloc culling, {num_lights} lights, {w}x{h}", flush=True)
        gs, os_ = _scene(num_lights, seed)
        f = H.camera_frustum((0, 0, 0), (1, 0, 0), (0, 0, 1), H.REFERENCE_FOV, w / h)
        ref_off, ref_list = gs.cull_lights(f, w, h)
        total = int(ref_off[-1])
        assert total == ref_list.size and total > 0
        offs, lists = _noalloc(gs, f, w, h, total)
        assert np.array_equal(offs, ref_off)
        assert np.array_equal(lists[:total], ref_list)
        o_off, o_list = os_.cull_lights(f, w, h)
        assert np.array_equal(offs, o_off) and np.array_equal(lists[:total], o_list)
        # truncated at capacity: the prefix of the full lists, nothing written past capacity
        cap = total // 3
        offs_t, lists_t = _noalloc(gs, f, w, h, cap)
        assert np.array_equal(offs_t, ref_off)
        assert np.array_equal(lists_t[:cap], ref_list[:cap])


def test_noalloc_culling_rejects_sizes_beyond_its_scratch():
    print("case: no-alloc culling beyond the scratch size is rejected before any launch", flush=True)
    gs, _ = _scene(4096, 3)
    f = H.camera_frustum((0, 0, 0), (1, 0, 0), (0, 0, 1), H.REFERENCE_FOV, 1.0)
    dev = G.dev()
    w = h = 8192  # 262 144 tiles x 129 words > 4 Mi words
    ntiles = (w // 16) * (h // 16)
    d_off, d_l = dev.alloc(4 * (ntiles + 1)), dev.alloc(64)
    st = dev.lib.rtq_cull_lights_noalloc(gs.handle, C.byref(f), w, h, d_off.ptr, d_l.ptr, 16, None)
    assert st == B.RTQ_ERR_LIMIT
