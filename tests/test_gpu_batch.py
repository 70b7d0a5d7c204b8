"""Batched form: one persistent launch over many scenes and uneven ray sets gives, item by item,
exactly what one plain launch per item gives (and what the oracle gives), with and without hipGraph
replay and for every autotuner setting."""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests import gpu_util as G
from tim_raytracing_amd import bindings as B
from tim_raytracing_amd import device as D
from tim_raytracing_amd import host as H

pytestmark = pytest.mark.gpu


def _scenes(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        tris = int(np.exp(rng.uniform(np.log(1000), np.log(40000))))
        b = H.BVHBuilder()
        b.generate(B.RTQ_SCENE_SPHERES, 1000 + k, max(1, tris // 256))
        b.build(16)
        blob, r = b.fill_gpu_buffer()
        out.append((b, blob, r, b.positions()))
    return out


def test_batch_matches_plain_and_oracle():
    print("case: 12 sphere scenes, 24 uneven ray sets (closest + any), batch vs plain vs oracle", flush=True)
    dev = G.dev()
    built = _scenes(12, 4)
    gscenes = [D.Scene(dev, blob, r, pos) for (_, blob, r, pos) in built]
    rng = np.random.default_rng(5)
    items, host_rays, bufs = [], [], []
    for k in range(24):
        si = int(rng.integers(0, len(gscenes)))
        n = int(np.exp(rng.uniform(np.log(100), np.log(60000))))
        b = built[si][0]
        if k % 2 == 0:
            f = H.camera_frustum((0, 0, 0), (1, 0, 0), (0, 0, 1), H.REFERENCE_FOV, 1.0)
            side = int(np.ceil(np.sqrt(n)))
            rays = H.frustum_rays(f, side, side, seed=k + 1)[:n].copy()
        else:
            rays = b.surface_rays(seed=k + 1, count=n)
        flags = B.RTQ_TRACE_CLOSEST if k % 3 else B.RTQ_TRACE_ANY
        d_r = dev.upload(rays)
        d_o = dev.alloc(16 * n)
        d_n = dev.alloc(4 * n) if flags == B.RTQ_TRACE_CLOSEST else None
        bufs.append((d_r, d_o, d_n))
        host_rays.append((si, flags, rays))
        items.append((si, flags, d_r.ptr, n, d_o.ptr, d_n.ptr if d_n else None))
    batch = D.Batch(dev, gscenes, items)

    def read_all():
        res = []
        for (si, flags, rays), (d_r, d_o, d_n) in zip(host_rays, bufs):
            n = rays.size
            if flags == B.RTQ_TRACE_CLOSEST:
                h = d_o.read(B.HIT_DTYPE, n)
                res.append((h.view(np.uint32).copy(), d_n.read(np.uint32, n)))
            else:
                res.append((d_o.read(np.uint32, n), None))
        return res

    batch.launch_plain()
    plain = read_all()
    # oracle on every item (closest: ids/t/u/v/nid bits; any: flags)
    for (si, flags, rays), (h, nid) in zip(host_rays, plain):
        _, blob, r, pos = built[si]
        sc = oracle.Scene(blob, r, pos)
        if flags == B.RTQ_TRACE_CLOSEST:
            o = sc.trace_closest(rays, threads=16)
            hv = h.reshape(-1, 4)
            assert np.array_equal(hv[:, 0], o["tri_id"])
            assert np.array_equal(hv[:, 1], o["t"].view(np.uint32))
            assert np.array_equal(hv[:, 2], o["u"].view(np.uint32))
            assert np.array_equal(hv[:, 3], o["v"].view(np.uint32))
            assert np.array_equal(nid, o["nid"])
        else:
            occ, _ = sc.trace_any(rays, threads=16)
            assert np.array_equal(h, occ)
    stream = dev.stream()
    sptr = stream.ptr
    for chunk, blocks, stack, graph, wait in ((64, 2, 24, 0, 0), (128, 4, 32, 0, 5), (512, 8, 64, 0, 64),
                                              (256, 6, 32, 1, 0), (1, 8, 24, 0, 0), (1, 2, 64, 0, 1),
                                              (1, 12, 32, 1, 48), (1, 8, 24, 0, 64)):
        batch.set_params(chunk_rays=chunk, blocks_per_cu=blocks, stack_depth=stack, use_graph=graph, leaf_wait=wait)
        for _ in range(3 if graph else 1):   # every graph replay must redo the work: outputs cleared before each
            for (d_r, d_o, d_n) in bufs:
                d_o.write(np.zeros(d_o.nbytes, np.uint8))
                if d_n is not None:
                    d_n.write(np.zeros(d_n.nbytes, np.uint8))
            batch.launch(sptr)
            stream.synchronize()
            got = read_all()
            for a, b in zip(plain, got):
                assert np.array_equal(a[0], b[0])
                if a[1] is not None:
                    assert np.array_equal(a[1], b[1])
    best, ms = batch.autotune(reps=1, stream=sptr)
    assert best.chunk_rays in (1, 64, 128, 256, 512) and best.leaf_wait <= 64 and ms > 0
    batch.launch(sptr)
    stream.synchronize()
    got = read_all()
    for a, b in zip(plain, got):
        assert np.array_equal(a[0], b[0])


def test_batch_rejects_bad_items():
    dev = G.dev()
    s = G.scene(1)["gpu"]
    buf = dev.alloc(1024)
    with pytest.raises(B.RtqError):
        D.Batch(dev, [s], [(3, B.RTQ_TRACE_CLOSEST, buf.ptr, 4, buf.ptr, None)])   # scene index
    with pytest.raises(B.RtqError):
        D.Batch(dev, [s], [(0, 9, buf.ptr, 4, buf.ptr, None)])                     # flags
    b = D.Batch(dev, [s], [(0, B.RTQ_TRACE_ANY, buf.ptr, 0, buf.ptr, None)])       # empty item is fine
    b.launch()
    with pytest.raises(B.RtqError):
        b.set_params(chunk_rays=100)


def test_batch_256_scenes_full_size():
    """The batched form at its stated size: 256 independently built scenes of LogU[1K, 2M] triangles,
    LogU[4K, 1M] rays per item (primary rays for even items, diffuse surface rays for odd ones), one
    autotuned, graph-replayed persistent launch. Every ray of every replay equals one plain launch per
    item, and the oracle agrees on 4096 strided rays of every item."""
    print("case: batch of 256 independent scenes (LogU[1K, 2M] tris), LogU[4K, 1M] rays per item, "
          "autotuned graph launch vs plain vs oracle samples", flush=True)
    builders, counts, pool = G.batch_pool()
    dev = G.dev()
    scenes = [D.Scene(dev, *p) for p in pool]
    items, bufs, host_rays = [], [], []
    for i in range(len(scenes)):
        n = int(counts[i])
        rays = H.batch_item_rays(builders[i], i, n)
        d_r, d_o, d_n = dev.upload(rays), dev.alloc(16 * n), dev.alloc(4 * n)
        bufs.append((d_r, d_o, d_n))
        host_rays.append(rays)
        items.append((i, B.RTQ_TRACE_CLOSEST, d_r.ptr, n, d_o.ptr, d_n.ptr))
    total = int(counts.sum())
    print(f"{total} rays in {len(items)} items", flush=True)
    batch = D.Batch(dev, scenes, items)
    stream = dev.stream()
    sptr = stream.ptr

    def read_all():
        return [(d_o.read(B.HIT_DTYPE, r.size).view(np.uint32).copy(), d_n.read(np.uint32, r.size))
                for (d_r, d_o, d_n), r in zip(bufs, host_rays)]

    batch.launch_plain(sptr)
    stream.synchronize()
    plain = read_all()
    best, ms = batch.autotune(reps=1, stream=sptr)
    print(f"autotuned: chunk {best.chunk_rays}, {best.blocks_per_cu} blocks/CU, stack {best.stack_depth}, "
          f"wait {best.leaf_wait}: {ms:.2f} ms per launch ({total / ms / 1e3:.0f} Mrays/s)", flush=True)
    batch.set_params(use_graph=1)
    for replay in range(2):
        for d_r, d_o, d_n in bufs:
            d_o.write(np.zeros(d_o.nbytes, np.uint8))
            d_n.write(np.zeros(d_n.nbytes, np.uint8))
        batch.launch(sptr)
        stream.synchronize()
        for k, (a, b) in enumerate(zip(plain, read_all())):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"replay {replay} item {k}"
    for i, ((h, nid), rays) in enumerate(zip(plain, host_rays)):
        idx = np.arange(0, rays.size, max(1, rays.size // 4096))[:4096]
        o = oracle.Scene(*pool[i]).trace_closest(np.ascontiguousarray(rays[idx]), threads=16)
        hv = h.reshape(-1, 4)[idx]
        assert np.array_equal(hv[:, 0], o["tri_id"]), f"item {i} tri_id"
        assert np.array_equal(hv[:, 1], o["t"].view(np.uint32)), f"item {i} t"
        assert np.array_equal(hv[:, 2], o["u"].view(np.uint32)) and np.array_equal(hv[:, 3], o["v"].view(np.uint32))
        assert np.array_equal(nid[idx], o["nid"]), f"item {i} nid"
