#!/usr/bin/env python3
"""Closest-hit ray throughput on one MI355X for the 1M-triangle sphere workload.

Workload: 4096 seeded UV spheres x 256 triangles = 1 048 576 fp32 triangles (centres U[-100,100]^3,
radius LogNormal(0, 0.7), random rotations), BVH from the reference builder restatement, 1920x1080
jittered primary rays from the origin looking +x (t limit 100), plus one any-hit shadow ray per hit
toward a point light at (0, 0, 150) (radius 1e4, so every hit traces).

A step is one closest-hit dispatch over all 2 073 600 primary rays with rays, BVH and hit buffers
resident in HBM. `value` is primary rays / step time. The shadow pass is timed separately and
reported beside it.

roofline: algorithmic bytes per ray B = 32 (ray) + 16 (hit) + 64*N_inner + 4*N_leaf + 48*N_tri, with the
per-ray counts from the counting variant of the same kernel (tests check those counts equal the
oracle's); achieved = sum(B) / average kernel time, peak = 8 TB/s HBM.
cpu_baseline: the scalar oracle (a restatement of the reference traversal) on a bounded sample of
the same rays, on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU time of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threads", type=int, default=0, help="host threads for build and CPU baseline")
    return ap.parse_args()


def host_threads(req: int) -> int:
    if req > 0:
        return req
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(blob, ranges, pos, rays, threads: int, target_s: float):
    import oracle  # the CPU baseline leg is the only place bench.py touches the oracle

    sc = oracle.Scene(blob, ranges, pos)
    probe = rays[:: max(1, rays.size // 20000)][:20000]
    t0 = time.perf_counter()
    sc.trace_closest(probe, threads=threads)
    dt = time.perf_counter() - t0
    rate = probe.size / max(dt, 1e-6)
    n = int(min(rays.size, max(probe.size, rate * target_s)))
    stride = max(1, rays.size // n)
    sample = np.ascontiguousarray(rays[::stride][:n])
    t0 = time.perf_counter()
    sc.trace_closest(sample, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(sample.size / dt / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample.size} of the {rays.size} primary rays (every {stride}th), closest hit, "
                  f"bitstack restatement of traverseBvh, {dt:.1f} s",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not part of "
                                   "this build", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from tim_raytracing_amd import bindings as B
    from tim_raytracing_amd import device as D
    from tim_raytracing_amd import host as H

    threads = host_threads(args.threads)
    t0 = time.perf_counter()
    builder, frustum, W, Hh = H.workload(2, scale=1.0, threads=threads)
    build_s = builder.stats().build_seconds
    blob, ranges = builder.fill_gpu_buffer()
    pos = builder.positions()
    rays = H.frustum_rays(frustum, W, Hh, seed=2)
    n = rays.size
    setup_s = time.perf_counter() - t0

    torch.cuda.init()
    dev = D.Device(0)
    scene = D.Scene(dev, blob, ranges, pos)
    info = scene.info()
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).cuda()
    d_hits = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    d_nid = torch.empty(n, dtype=torch.int32, device="cuda")
    d_masks = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    def closest():
        scene.dispatch(d_rays.data_ptr(), n, B.RTQ_TRACE_CLOSEST, d_hits.data_ptr(), d_nid.data_ptr(), sptr)

    # per-ray visit counts for the algorithmic bytes (counting variant, untimed)
    d_cnt = torch.empty(n * 3, dtype=torch.int32, device="cuda")
    scene.dispatch_counted(d_rays.data_ptr(), n, B.RTQ_TRACE_CLOSEST, d_hits.data_ptr(), d_nid.data_ptr(),
                           d_cnt.data_ptr(), sptr)
    torch.cuda.synchronize()
    cnt = d_cnt.cpu().numpy().astype(np.int64).reshape(n, 3)
    sum_inner, sum_leaf, sum_tri = (int(x) for x in cnt.sum(0))
    alg_bytes = 48 * n + 64 * sum_inner + 4 * sum_leaf + 48 * sum_tri
    del d_cnt

    for _ in range(args.warmup):
        closest()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        closest()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps

    hits = d_hits.cpu().numpy().view(B.HIT_DTYPE)
    hit_mask = hits["tri_id"] != B.RTQ_NO_HIT
    n_hit = int(hit_mask.sum())

    # shadow pass: one any-hit ray per hit toward the point light (light 0 of every leaf list)
    zero_sun = np.zeros(3, np.float32)
    for _ in range(max(1, args.warmup // 2)):
        scene.shadow_masks(d_rays.data_ptr(), d_hits.data_ptr(), d_nid.data_ptr(), n, zero_sun, 0,
                           d_masks.data_ptr(), sptr)
    ev0.record(stream)
    for _ in range(args.steps):
        scene.shadow_masks(d_rays.data_ptr(), d_hits.data_ptr(), d_nid.data_ptr(), n, zero_sun, 0,
                           d_masks.data_ptr(), sptr)
    ev1.record(stream)
    torch.cuda.synchronize()
    shadow_ms = ev0.elapsed_time(ev1) / args.steps
    masks = d_masks.cpu().numpy()

    achieved = alg_bytes / (ms * 1e-3) / 1e9
    result = {
        "metric": "closest-hit Mrays/s @1M tris (frac of HBM roofline)",
        "value": round(n / (ms * 1e-3) / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "spheres-1M-primary-1920x1080",
            "triangles": int(builder.stats().num_triangles),
            "rays_per_step": n,
            "tmax": 100.0,
            "bvh_nodes": int(info.num_nodes),
            "leaf_triangle_slots": int(info.num_leaf_slots),
            "bvh_max_depth": int(info.max_depth),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "kTraceClosest<32,false>",
            "algorithmic_bytes_per_launch": int(alg_bytes),
            "mean_bytes_per_ray": round(alg_bytes / n, 1),
            "mean_inner_nodes": round(sum_inner / n, 2),
            "mean_leaves": round(sum_leaf / n, 3),
            "mean_triangle_tests": round(sum_tri / n, 2),
            "roofline_mrays_s": round(HBM_PEAK_GBS * 1e9 / (alg_bytes / n) / 1e6, 1),
        },
        "hit_rate": round(n_hit / n, 4),
        "shadow": {
            "rays": n_hit,
            "ms_per_step": round(shadow_ms, 4),
            "anyhit_mrays_s": round(n_hit / (shadow_ms * 1e-3) / 1e6, 2),
            "occluded_frac": round(float(((masks >> 1) & 1)[hit_mask].mean()) if n_hit else 0.0, 4),
        },
        "host": {"threads": threads, "bvh_build_s": round(build_s, 2), "setup_s": round(setup_s, 2)},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(blob, ranges, pos, rays, threads, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
