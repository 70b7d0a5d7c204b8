"""On-disk scene cache (include/rtq_build.h, rtq_cache_*): round trip, keying and corruption checks."""
import os

import numpy as np
import pytest

from tim_raytracing_amd import bindings as B
from tim_raytracing_amd import host as H


def _builder(seed=3, spheres=40, min_obj=4):
    b = H.BVHBuilder()
    b.generate(B.RTQ_SCENE_SPHERES, seed, spheres)
    b.add_sphere_light((1, 2, 3), 50.0)
    b.set_parameters(min_obj_per_node=min_obj)
    b.build(2)
    return b


def test_cache_round_trip_is_exact(tmp_path):
    b = _builder()
    path = str(tmp_path / "scene.rtqbvh")
    b.save_cache(path)
    loaded = H.load_cache(path, b.input_hash())
    assert loaded is not None
    blob, r, pos, key = loaded
    blob0, r0 = b.fill_gpu_buffer()
    assert key == b.input_hash()
    assert np.array_equal(blob, blob0)
    for name in ("triangles", "primitives", "materials", "lights", "nodes", "leaf_data", "blas"):
        assert (getattr(r, name).offset, getattr(r, name).size) == (getattr(r0, name).offset, getattr(r0, name).size)
    assert np.array_equal(pos.view(np.uint32), b.positions().reshape(-1, 3).view(np.uint32))
    # expected hash 0 accepts any key
    assert H.load_cache(path, 0)[3] == key
    assert not os.path.exists(path + ".tmp")


def test_cache_key_follows_inputs(tmp_path):
    a, same = _builder(), _builder()
    assert a.input_hash() == same.input_hash()
    assert _builder(seed=4).input_hash() != a.input_hash()            # other geometry
    assert _builder(min_obj=8).input_hash() != a.input_hash()         # other build parameters
    c = H.BVHBuilder()
    c.generate(B.RTQ_SCENE_SPHERES, 3, 40)
    c.set_parameters(min_obj_per_node=4)
    c.build(1)
    assert c.input_hash() != a.input_hash()                           # same geometry, no light
    path = str(tmp_path / "a.rtqbvh")
    a.save_cache(path)
    assert H.load_cache(path, c.input_hash()) is None                 # stale: a miss, not an error
    assert H.load_cache(str(tmp_path / "missing.rtqbvh"), 0) is None


def test_cache_rejects_corrupt_files(tmp_path):
    b = _builder()
    path = str(tmp_path / "s.rtqbvh")
    b.save_cache(path)
    raw = bytearray(open(path, "rb").read())
    flipped = bytearray(raw)
    flipped[len(raw) // 2] ^= 0x40                                    # one bit in the payload
    for bad in (flipped, raw[:-5], b"XYZ" + raw[3:], raw[:50]):
        p = str(tmp_path / "bad.rtqbvh")
        open(p, "wb").write(bytes(bad))
        with pytest.raises(B.RtqError):
            H.load_cache(p, 0)


def test_cache_save_needs_a_built_builder(tmp_path):
    b = H.BVHBuilder()
    b.generate(B.RTQ_SCENE_SPHERES, 1, 4)
    with pytest.raises(B.RtqError):
        b.save_cache(str(tmp_path / "x.rtqbvh"))
