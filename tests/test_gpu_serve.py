"""Served form: results identical alone and in the mix, admission under the HBM budget,
queue-full backpressure and cancellation."""
import threading
import time

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
        tris = int(np.exp(rng.uniform(np.log(2000), np.log(60000))))
        b = H.BVHBuilder()
        b.generate(B.RTQ_SCENE_SPHERES, 2000 + k, max(1, tris // 256))
        b.build(16)
        out.append(b)
    return out


def _requests(builders, count, seed, lo=200, hi=80000):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        si = int(rng.integers(0, len(builders)))
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        rays = builders[si].surface_rays(seed=10 + k, count=n)
        flags = B.RTQ_TRACE_CLOSEST if rng.random() < 0.5 else B.RTQ_TRACE_ANY
        reqs.append((si, flags, rays))
    return reqs


def test_serve_mix_equals_alone():
    print("case: 8 resident scenes, 40 requests from 4 client threads, results vs alone", flush=True)
    dev = G.dev()
    builders = _scenes(8, 7)
    scenes = [D.Scene.from_builder(dev, b) for b in builders]
    reqs = _requests(builders, 40, 8)
    # alone: one dispatch per request, each checked against the oracle on every ray
    orc = [oracle.Scene(*b.fill_gpu_buffer(), b.positions()) for b in builders]
    alone = []
    for si, flags, rays in reqs:
        if flags == B.RTQ_TRACE_CLOSEST:
            g = scenes[si].trace_closest(rays)
            G.assert_closest_equal(g, orc[si].trace_closest(rays, threads=16), "served request alone vs oracle")
            alone.append((g["hits"].view(np.uint32).copy(), g["nid"]))
        else:
            occ = scenes[si].trace_any(rays)
            assert np.array_equal(occ, orc[si].trace_any(rays, threads=16)[0])
            alone.append((occ, None))
    # mixed: a budget that holds only part of the stream at once, 4 submitting threads; chunked and
    # lane-refilling batch kernels
    for params in (None, B.BatchParams(B.RTQ_BATCH_REFILL, 8, 64, 0, 64)):
        srv = D.Server(dev, scenes, budget_bytes=8 << 20, params=params)
        handles = [None] * len(reqs)

        def client(t):
            for k in range(t, len(reqs), 4):
                si, flags, rays = reqs[k]
                handles[k] = srv.submit(si, flags, rays)

        th = [threading.Thread(target=client, args=(t,)) for t in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        batches = set()
        for k, h in enumerate(handles):
            assert h.wait() == B.RTQ_OK
            st = h.stats()
            batches.add(st.batch_id)
            assert st.latency_ms >= st.queued_ms >= 0
            a_out, a_nid = alone[k]
            if reqs[k][1] == B.RTQ_TRACE_CLOSEST:
                assert np.array_equal(h.out.view(np.uint32), a_out)
                assert np.array_equal(h.nid, a_nid)
            else:
                assert np.array_equal(h.out, a_out)
            h.release()
        assert len(batches) < len(reqs)  # requests were batched together
        srv.close()
    # plain comparison mode gives the same results
    srv = D.Server(dev, scenes, budget_bytes=8 << 20, plain=True)
    hs = [srv.submit(si, flags, rays) for si, flags, rays in reqs[:10]]
    for k, h in enumerate(hs):
        assert h.wait() == B.RTQ_OK
        if reqs[k][1] == B.RTQ_TRACE_CLOSEST:
            assert np.array_equal(h.out.view(np.uint32), alone[k][0])
        else:
            assert np.array_equal(h.out, alone[k][0])
        h.release()
    srv.close()


def test_serve_device_requests_equal_alone():
    """Requests whose rays and results live in device memory (rtq_server_submit_device) are traced in
    place: mixed with host-buffer requests, batched, from several threads, each equals the request
    traced alone and the oracle; misaligned and undersized device buffers are rejected."""
    print("case: 8 resident scenes, 40 device-resident requests + 10 host requests, 4 threads, vs alone", flush=True)
    dev = G.dev()
    builders = _scenes(8, 17)
    scenes = [D.Scene.from_builder(dev, b) for b in builders]
    reqs = _requests(builders, 50, 18)
    orc = [oracle.Scene(*b.fill_gpu_buffer(), b.positions()) for b in builders]
    alone = []
    for si, flags, rays in reqs:
        if flags == B.RTQ_TRACE_CLOSEST:
            g = scenes[si].trace_closest(rays)
            G.assert_closest_equal(g, orc[si].trace_closest(rays, threads=16), "device request alone vs oracle")
            alone.append((g["hits"].view(np.uint32).copy(), g["nid"]))
        else:
            occ = scenes[si].trace_any(rays)
            assert np.array_equal(occ, orc[si].trace_any(rays, threads=16)[0])
            alone.append((occ, None))
    bufs = []
    for si, flags, rays in reqs:
        n = rays.size
        closest = flags == B.RTQ_TRACE_CLOSEST
        bufs.append((dev.upload(rays), dev.alloc((16 if closest else 4) * n), dev.alloc(4 * n) if closest else None))
    for params in (None, B.BatchParams(64, 4, 64, 0, 0)):
        srv = D.Server(dev, scenes, budget_bytes=4 << 20, params=params)
        handles = [None] * len(reqs)

        def client(t):
            for k in range(t, len(reqs), 4):
                si, flags, rays = reqs[k]
                if k % 5 == 4:   # every fifth request through host buffers, in the same batches
                    handles[k] = srv.submit(si, flags, rays)
                else:
                    d_r, d_o, d_n = bufs[k]
                    d_o.write(np.zeros(d_o.nbytes, np.uint8))
                    handles[k] = srv.submit_device(si, flags, d_r, rays.size, d_o, d_n)

        th = [threading.Thread(target=client, args=(t,)) for t in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        batches = set()
        for k, h in enumerate(handles):
            assert h.wait() == B.RTQ_OK
            batches.add(h.stats().batch_id)
            si, flags, rays = reqs[k]
            a_out, a_nid = alone[k]
            if k % 5 == 4:
                got = h.out.view(np.uint32) if flags == B.RTQ_TRACE_CLOSEST else h.out
                assert np.array_equal(got, a_out)
            elif flags == B.RTQ_TRACE_CLOSEST:
                _, d_o, d_n = bufs[k]
                assert np.array_equal(d_o.read(np.uint32, 4 * rays.size), a_out)
                assert np.array_equal(d_n.read(np.uint32, rays.size), a_nid)
            else:
                assert np.array_equal(bufs[k][1].read(np.uint32, rays.size), a_out)
            h.release()
        assert len(batches) < len(reqs)
        srv.close()
    # plain comparison mode on device requests
    srv = D.Server(dev, scenes, budget_bytes=4 << 20, plain=True)
    for k in range(6):
        si, flags, rays = reqs[k]
        d_r, d_o, d_n = bufs[k]
        d_o.write(np.zeros(d_o.nbytes, np.uint8))
        h = srv.submit_device(si, flags, d_r, rays.size, d_o, d_n)
        assert h.wait() == B.RTQ_OK
        got = d_o.read(np.uint32, (4 if flags == B.RTQ_TRACE_CLOSEST else 1) * rays.size)
        assert np.array_equal(got, alone[k][0])
        h.release()
    # misaligned rays and a too-small output are rejected before anything is queued
    d_r, d_o, _ = bufs[1]
    with pytest.raises(B.RtqError):
        srv.submit_device(0, B.RTQ_TRACE_ANY, D.DeviceView(d_r, 8, 32 * 10), 10, d_o)
    with pytest.raises(ValueError):
        srv.submit_device(0, B.RTQ_TRACE_CLOSEST, d_r, reqs[1][2].size, D.DeviceView(d_o, 0, 16))
    # the library checks the buffers itself: host memory, and a range running past its allocation's end
    import ctypes as C
    lib, h = dev.lib, C.c_void_p()
    hb = dev.host_alloc(4096)
    small = dev.alloc(4096)   # 128 rays
    big_out = dev.alloc(4 << 20)
    for rays_ptr, out_ptr, n in ((hb.ptr, big_out.ptr, 16), (small.ptr, hb.ptr, 16), (small.ptr, big_out.ptr, 100_000),
                                 (small.ptr + 32 * 60, big_out.ptr, 100_000)):
        st = lib.rtq_server_submit_device(srv.handle, 0, B.RTQ_TRACE_ANY, rays_ptr, n, out_ptr, None, C.byref(h))
        assert st == 1, (rays_ptr, out_ptr, n, st)  # RTQ_ERR_INVALID_ARGUMENT, nothing queued
    # the allocation's 128 rays exactly fit
    small.write(reqs[1][2][:128])
    req = srv.submit_device(0, B.RTQ_TRACE_ANY, small, 128, D.DeviceView(big_out, 0, 4 * 128))
    assert req.wait() == B.RTQ_OK
    req.release()
    srv.close()


def test_serve_admission_limits():
    print("case: server budget rejection, queue-full, cancellation", flush=True)
    dev = G.dev()
    b = _scenes(1, 9)[0]
    sc = D.Scene.from_builder(dev, b)
    srv = D.Server(dev, [sc], budget_bytes=1 << 20, max_queue=3)
    big = b.surface_rays(seed=1, count=40000)        # 52 B/ray > 1 MiB budget
    with pytest.raises(B.RtqError) as e:
        srv.submit(0, B.RTQ_TRACE_CLOSEST, big)
    assert e.value.status == 7  # RTQ_ERR_BUDGET
    small = b.surface_rays(seed=2, count=15000)       # ~780 KB: one fits at a time
    hs = []
    full = False
    for k in range(40):
        try:
            hs.append(srv.submit(0, B.RTQ_TRACE_CLOSEST, small))
        except B.RtqError as err:
            assert err.status == 8  # RTQ_ERR_QUEUE_FULL
            full = True
            break
    assert full
    # cancel whatever is still queued; the rest completes
    cancelled = sum(1 for h in hs if h.cancel() == B.RTQ_OK)
    for h in hs:
        st = h.wait()
        assert st in (B.RTQ_OK, 9)
        h.release()
    assert cancelled >= 1
    srv.close()


def test_serve_requests_outlive_server():
    """Destroying the server with requests queued and in flight: admitted ones complete with correct
    results, queued ones are cancelled, and every handle can still be waited on, queried and released
    afterwards."""
    print("case: server destroyed before its request handles are released", flush=True)
    dev = G.dev()
    b = _scenes(1, 11)[0]
    sc = D.Scene.from_builder(dev, b)
    srv = D.Server(dev, [sc], budget_bytes=4 << 20, max_inflight=1)
    rays = [b.surface_rays(seed=40 + k, count=30000) for k in range(12)]
    hs = [srv.submit(0, B.RTQ_TRACE_ANY, r) for r in rays]
    srv.close()
    done = 0
    for h, r in zip(hs, rays):
        st = h.wait()
        assert st in (B.RTQ_OK, 9)  # 9: RTQ_ERR_CANCELLED
        if st == B.RTQ_OK:
            assert np.array_equal(h.out, sc.trace_any(r))
            assert h.stats().latency_ms > 0
            done += 1
        assert h.cancel() != B.RTQ_OK  # finished: nothing left to cancel
        h.release()
    assert done >= 1


def test_serve_64_scenes_poisson_full_size():
    """The served form at its stated size: 64 resident scenes of LogU[10K, 3M] triangles and 240
    Poisson-timed requests of LogU[1K, 2M] rays (half closest, half any hit) submitted by 4 client
    threads under a 2 GiB budget. Every ray of every served request equals the request traced alone (one
    plain dispatch), also when the same requests come as device-buffer requests, and the oracle agrees on
    4096 strided rays of every request."""
    print("case: 64 resident scenes (LogU[10K, 3M] tris), 240 Poisson requests of LogU[1K, 2M] rays from "
          "4 threads, served vs alone vs oracle samples", flush=True)
    nreq = 240
    builders, rng, sizes, req_scene, req_any, pool = G.serve_pool(nreq)
    dev = G.dev()
    scenes = [D.Scene(dev, *p) for p in pool]
    reqs = []
    for k in range(nreq):
        si = int(req_scene[k])
        rays = builders[si].surface_rays(seed=100 + k, count=int(sizes[k]))
        reqs.append((si, B.RTQ_TRACE_ANY if req_any[k] else B.RTQ_TRACE_CLOSEST, rays))
    print(f"{int(sizes.sum())} rays in {nreq} requests", flush=True)
    alone = []
    for si, flags, rays in reqs:
        if flags == B.RTQ_TRACE_CLOSEST:
            g = scenes[si].trace_closest(rays)
            alone.append((g["hits"].view(np.uint32).copy(), g["nid"]))
        else:
            alone.append((scenes[si].trace_any(rays), None))
    # Poisson arrivals at 400 requests/s, spread over 4 submitting threads
    arrivals = np.cumsum(rng.exponential(1.0 / 400.0, nreq))
    srv = D.Server(dev, scenes, budget_bytes=2 << 30)
    handles = [None] * nreq
    start = time.perf_counter()

    def client(t):
        for k in range(t, nreq, 4):
            wait = start + arrivals[k] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            si, flags, rays = reqs[k]
            handles[k] = srv.submit(si, flags, rays)

    th = [threading.Thread(target=client, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    lat, batches = [], set()
    for k, h in enumerate(handles):
        assert h.wait() == B.RTQ_OK
        st = h.stats()
        lat.append(st.latency_ms)
        batches.add(st.batch_id)
        a_out, a_nid = alone[k]
        if reqs[k][1] == B.RTQ_TRACE_CLOSEST:
            assert np.array_equal(h.out.view(np.uint32), a_out), f"request {k}"
            assert np.array_equal(h.nid, a_nid), f"request {k} nid"
        else:
            assert np.array_equal(h.out, a_out), f"request {k}"
    print(f"{len(batches)} launches for {nreq} requests, p50 {np.percentile(lat, 50):.1f} ms, "
          f"p99 {np.percentile(lat, 99):.1f} ms", flush=True)
    assert len(batches) < nreq
    for h in handles:
        h.release()
    # the same stream as device-buffer requests (rays and results in HBM, traced in place), all submitted
    # at once from 4 threads onto the same server: every ray again equals the request traced alone
    bufs = []
    for si, flags, rays in reqs:
        closest = flags == B.RTQ_TRACE_CLOSEST
        bufs.append((dev.upload(rays), dev.alloc((16 if closest else 4) * rays.size),
                     dev.alloc(4 * rays.size) if closest else None))
    handles = [None] * nreq

    def dclient(t):
        for k in range(t, nreq, 4):
            si, flags, rays = reqs[k]
            d_r, d_o, d_n = bufs[k]
            handles[k] = srv.submit_device(si, flags, d_r, rays.size, d_o, d_n)

    th = [threading.Thread(target=dclient, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for k, h in enumerate(handles):
        assert h.wait() == B.RTQ_OK
        si, flags, rays = reqs[k]
        _, d_o, d_n = bufs[k]
        a_out, a_nid = alone[k]
        if flags == B.RTQ_TRACE_CLOSEST:
            assert np.array_equal(d_o.read(np.uint32, 4 * rays.size), a_out), f"device request {k}"
            assert np.array_equal(d_n.read(np.uint32, rays.size), a_nid), f"device request {k} nid"
        else:
            assert np.array_equal(d_o.read(np.uint32, rays.size), a_out), f"device request {k}"
        h.release()
    print(f"{nreq} device-buffer requests equal alone on every ray", flush=True)
    del bufs
    srv.close()
    # the oracle on strided samples of every request, one call per scene and mode
    by_scene = {}
    for k, (si, flags, rays) in enumerate(reqs):
        idx = np.arange(0, rays.size, max(1, rays.size // 4096))[:4096]
        by_scene.setdefault((si, flags), []).append((k, idx))
    for (si, flags), parts in by_scene.items():
        sample = np.concatenate([reqs[k][2][idx] for k, idx in parts])
        sc = oracle.Scene(*pool[si])
        if flags == B.RTQ_TRACE_CLOSEST:
            o = sc.trace_closest(sample, threads=16)
            got = np.concatenate([alone[k][0].reshape(-1, 4)[idx] for k, idx in parts])
            nid = np.concatenate([alone[k][1][idx] for k, idx in parts])
            assert np.array_equal(got[:, 0], o["tri_id"]) and np.array_equal(got[:, 1], o["t"].view(np.uint32))
            assert np.array_equal(got[:, 2], o["u"].view(np.uint32)) and np.array_equal(got[:, 3], o["v"].view(np.uint32))
            assert np.array_equal(nid, o["nid"])
        else:
            got = np.concatenate([alone[k][0][idx] for k, idx in parts])
            assert np.array_equal(got, sc.trace_any(sample, threads=16)[0])
