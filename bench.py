#!/usr/bin/env python3
"""Closest-hit ray throughput on one MI355X for the 1M-triangle sphere workload.

Workload: 4096 seeded UV spheres x 256 triangles = 1 048 576 fp32 triangles (centres U[-100,100]^3,
radius LogNormal(0, 0.7), random rotations), BVH from the reference builder restatement, 1920x1080
jittered primary rays from the origin looking +x (t limit 100), plus one any-hit shadow ray per hit
toward a point light at (0, 0, 150) (radius 1e4, so every hit traces).

A step is one closest-hit dispatch over all 2 073 600 primary rays with rays, BVH and hit buffers
resident in HBM. `value` is primary rays / step time. The shadow pass is timed separately and
reported beside it.

roofline: the coherent traversal kernels are held by VALU instruction issue (the L2 and Infinity Cache
serve nearly all the bytes the rays read): achieved = the kernel's SQ_INSTS_VALU per launch (committed
PMC summary, profiles/pmc_<workload>.json) / the launch time measured here with HIP events, peak =
614.4 G wave64 VALU instructions/s; traffic = bytes beyond the L2 per launch (request-size counters
TCC_EA0_RDREQ_*/WRREQ_* weighted by request size) and hbm_frac their share of 8 TB/s, an upper bound
of the HBM load since the Infinity Cache serves part of them. Whichever fraction is higher names the
bound. The reference-layout algorithmic bytes per ray,
B = 32 (ray) + 16 (hit) + 64*N_inner + 4*N_leaf + 48*N_tri, with per-ray counts from the counting
variant of the same kernel (tests check those counts equal the oracle's), are reported beside it.
cpu_baseline: the scalar oracle (a restatement of the reference traversal) on a bounded sample of
the same rays, on the host cores. The same leg checks every ray of the timed launch against the
oracle (parity) and states the FP64 error of its hits (fp64_error).
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
    ap.add_argument("--mode", choices=("primary", "incoherent", "batch", "serve", "tlas", "copy"), default="primary",
                    help="primary: the headline workload; incoherent: 3 Mi-tri soup, 2 diffuse bounces + tile "
                         "culling; batch: many scenes in one persistent launch vs one launch each; serve: "
                         "Poisson request stream, continuous batching vs one launch per request; tlas: any hit "
                         "over a two-level scene of 1024 BLAS instances")
    ap.add_argument("--scale", type=float, default=1.0, help="triangle-count scale for incoherent/batch/serve")
    ap.add_argument("--distinct", type=int, default=0,
                    help="batch/serve: number of distinct BVHs built; scene slots reuse them round-robin "
                         "(0 = one build per slot)")
    ap.add_argument("--requests", type=int, default=1000, help="serve: requests in the stream")
    ap.add_argument("--cache-dir", default=None, help="batch/serve: on-disk scene cache directory (built scenes are "
                                                      "stored there and loaded by later runs)")
    ap.add_argument("--max-batch-rays", type=int, default=4_000_000,
                    help="serve: rays per served launch (0 = no cap)")
    ap.add_argument("--chunk-rays", type=int, default=1, help="serve: rays a wave takes per grab (1 = lane refill)")
    ap.add_argument("--max-inflight", type=int, default=0, help="serve: launches on the device at once (0 = "
                                                                "library default)")
    ap.add_argument("--ray-order", choices=("row", "tile"), default="tile",
                    help="primary: order of the ray buffer (tile: 8x8 pixel tiles, one per wave, as a 2D "
                         "compute dispatch of the camera pass covers the frame; row: y*W + x)")
    ap.add_argument("--dispatch", choices=("auto", "persistent", "range", "grid"), default="auto",
                    help="persistent: lanes refill from all rays; range: each wave refills from its own "
                         "contiguous range; grid: one lane per ray; auto: grid for the coherent primary "
                         "rays, persistent for incoherent bounces")
    ap.add_argument("--leaf-wait", type=int, default=0,
                    help="lanes ready for leaf tests that end a round of descent, 1..64 (0 = library default)")
    ap.add_argument("--sort", type=int, choices=(-1, 0, 1), default=-1,
                    help="primary/incoherent: sort rays by direction octant and origin cell before tracing, inside "
                         "the timed dispatch (-1 = auto: on for incoherent, off for primary)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="persistent dispatch: blocks per CU (0 = library "
                                                                 "default)")
    ap.add_argument("--batch-params", default=None,
                    help="batch: chunk_rays,blocks_per_cu,stack_depth,leaf_wait to use instead of autotuning "
                         "(to profile the configuration an earlier run picked)")
    return ap.parse_args()


def tile_order(width: int, height: int, t: int = 8) -> np.ndarray:
    """Pixel indices y*width + x listed tile by tile (t x t tiles, tiles row-major, pixels row-major
    inside a tile): one wave of 64 lanes then traces one 8x8 tile, as a 2D compute dispatch does."""
    assert width % t == 0 and height % t == 0
    y, x = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    idx = (y * width + x).reshape(height // t, t, width // t, t).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(idx.reshape(-1))


def stack_class(depth: int) -> int:
    """Stack depth parameter of the batched and served forms (24, 32 or 64)."""
    return 24 if depth <= 24 else (32 if depth <= 32 else 64)


def kernel_stack(depth: int) -> int:
    """LDS stack entries the single-scene kernels are instantiated with for a tree of this depth."""
    return 22 if depth <= 22 else stack_class(depth)


def configure_dispatch(dev, args, auto: str, auto_sort: int = 0):
    """Apply --dispatch / --sort / --leaf-wait / --blocks-per-cu to the device; returns a label for the
    output."""
    from tim_raytracing_amd import bindings as B

    if args.dispatch == "auto":
        args.dispatch = auto
    if args.sort < 0:
        args.sort = auto_sort
    dev.set_option(B.RTQ_OPTION_PERSISTENT, {"grid": 0, "persistent": 1, "range": 2}[args.dispatch])
    dev.set_option(B.RTQ_OPTION_SORT_RAYS, args.sort)
    dev.set_option(B.RTQ_OPTION_LEAF_WAIT, args.leaf_wait)
    if args.blocks_per_cu:
        dev.set_option(B.RTQ_OPTION_BLOCKS_PER_CU, args.blocks_per_cu)
    return {"mode": args.dispatch, "blocks_per_cu": dev.get_option(B.RTQ_OPTION_BLOCKS_PER_CU) or "auto (7; 12 two-level)",
            "sort_rays": args.sort,
            "leaf_wait": args.leaf_wait}


def kernel_name(args, depth: int, any_hit: bool = False) -> str:
    if args.dispatch in ("persistent", "range"):
        base = "kTracePersistent" if args.dispatch == "persistent" else "kTraceWaveRange"
        return f"{base}<{kernel_stack(depth)},{'true' if any_hit else 'false'}>"
    return f"{'kTraceAny' if any_hit else 'kTraceClosest'}<{kernel_stack(depth)},false>"


# Published MI355X figures: 8 TB/s HBM3E; 256 CUs at 2.4 GHz peak engine clock. A wave64 VALU
# instruction occupies a 16-lane SIMD for 4 cycles and a CU has 4 SIMDs, so the chip issues at most
# 256 x 2.4e9 = 614.4 G wave64 VALU instructions per second (the published 157.3 TFLOP/s fp32 figure
# / 2 for packed / 2 for FMA / 64 lanes is the same number).
VALU_PEAK_GINSTR = 256 * 2.4


def pmc_record(workload: str):
    """The committed PMC summary of a workload (profiles/pmc_<workload>.json), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{workload}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        return json.load(fh), os.path.relpath(path, ROOT)


def measured_traffic(workload: str, kernel: str):
    """Bytes beyond the L2 per launch (what the L2 reads from and writes to the Infinity Cache / HBM)
    from the committed PMC summary of this workload and kernel, if present: the request-size counters
    (TCC_EA0_RDREQ_{32B,64B,128B} and TCC_EA0_WRREQ_64B weighted by their sizes, which read a 1 GiB copy
    as 1.000 GiB each way), else FETCH_SIZE / WRITE_SIZE scaled by what they read for that copy."""
    rec, src = pmc_record(workload)
    k = (rec or {}).get("kernels", {}).get(kernel)
    if not k:
        return None, None
    if "rd_bytes" in k:
        return int(k["rd_bytes"] + k["wr_bytes"]), src
    if "fetch_kib" not in k:
        return None, None
    b = 1024.0 * (k["fetch_kib"] * rec["fetch_scale"] + k["write_kib"] * rec["write_scale"])
    return int(b), src


def roofline(workload: str, kernel: str, ms: float, algorithmic: dict | None = None) -> dict:
    """The kernel's position under its two roofs, from the committed PMC summary of the same workload
    and kernel and the launch time measured here: VALU issue (SQ_INSTS_VALU per launch / launch time
    against 614.4 G wave64 instructions/s) and HBM (traffic = HBM bytes per launch from FETCH_SIZE /
    WRITE_SIZE, / launch time against 8 TB/s). bound/achieved/peak/frac name the roof with the higher
    fraction. `algorithmic` (the reference-layout bytes per launch) is informational: the caches serve
    most of those bytes, so they are not a fraction of any roof."""
    rec, src = pmc_record(workload)
    k = (rec or {}).get("kernels", {}).get(kernel) or {}
    valu = k.get("sq_insts_valu")
    traffic, _ = measured_traffic(workload, kernel)
    sec = ms * 1e-3
    valu_rate = valu / sec / 1e9 if valu else None
    hbm_rate = traffic / sec / 1e9 if traffic else None
    valu_frac = valu_rate / VALU_PEAK_GINSTR if valu_rate else None
    hbm_frac = hbm_rate / HBM_PEAK_GBS if hbm_rate else None
    use_hbm = hbm_frac is not None and (valu_frac is None or hbm_frac > valu_frac)
    out = {
        "bound": "hbm" if use_hbm else "valu_issue",
        "achieved": round(hbm_rate if use_hbm else valu_rate, 1) if (hbm_rate if use_hbm else valu_rate) else None,
        "peak": HBM_PEAK_GBS if use_hbm else VALU_PEAK_GINSTR,
        "unit": "GB/s" if use_hbm else "G wave64 VALU instr/s",
        "frac": round(hbm_frac if use_hbm else valu_frac, 4) if (hbm_frac if use_hbm else valu_frac) else None,
        "traffic": traffic,
        "valu_frac": round(valu_frac, 4) if valu_frac else None,
        "hbm_frac": round(hbm_frac, 4) if hbm_frac else None,
        "valu_instructions_per_launch": valu,
        "valu_g_per_s": round(valu_rate, 1) if valu_rate else None,
        "hbm_gb_s": round(hbm_rate, 1) if hbm_rate else None,
        "active_lanes_per_valu_instr": k.get("active_lane_frac"),
        "counter_source": src if k else None,
        "kernel": kernel,
    }
    if algorithmic is not None:
        out["algorithmic"] = algorithmic
    return out


def host_threads(req: int) -> int:
    if req > 0:
        return req
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_oracle(fn, rays, threads: int, target_s: float, what: str):
    """Repeats fn(sample, threads) over a strided sample of `rays` sized for about target_s of wall
    time; -> (Mrays/s, description)."""
    probe = rays[:: max(1, rays.size // 20000)][:20000]
    t0 = time.perf_counter()
    fn(np.ascontiguousarray(probe), threads)
    rate = probe.size / max(time.perf_counter() - t0, 1e-6)
    n = int(min(rays.size, max(probe.size, rate * target_s)))
    stride = max(1, rays.size // n)
    sample = np.ascontiguousarray(rays[::stride][:n])
    passes, traced = 0, 0
    t0 = time.perf_counter()
    while True:
        fn(sample, threads)
        passes += 1
        traced += sample.size
        dt = time.perf_counter() - t0
        if dt >= target_s or passes >= 1000:
            break
    return traced / dt / 1e6, (f"{sample.size} of the {rays.size} {what} (stride {stride}) x {passes} passes, "
                               f"{threads} threads, {dt:.1f} s")


def check_against_oracle(sc, rays, g_hits, g_nid, g_masks, threads: int) -> tuple[dict, dict]:
    """Every ray of the timed launch against the oracle (a restatement of the reference traversal):
    closest-hit triangle id, leaf node id and the bits of t, u, v; the shadow masks of the timed
    shadow pass; and the FP64 error of the fp32 hits (t relative to the FP64 intersection of the same
    triangle, after the reference's t - 1e-5 shrink; u, v absolute)."""
    o = sc.trace_closest(rays, threads=threads)
    bad = (g_hits["tri_id"] != o["tri_id"]) | (g_nid != o["nid"])
    for k in ("t", "u", "v"):
        bad |= g_hits[k].view(np.uint32) != o[k].view(np.uint32)
    om = sc.shadow_masks(rays, o["t"], o["nid"], np.zeros(3, np.float32), 0, threads=threads)
    hit = o["tri_id"] != 0xFFFFFFFF
    parity = {"rays_checked": int(rays.size), "mismatches": int(bad.sum()),
              "shadow_checked": int(rays.size), "shadow_mismatches": int((g_masks != om).sum()),
              "hits": int(hit.sum()), "near_tie_exempted": 0,
              "checker": "oracle/rtq_oracle.c bitstack traverseBvh + computeShadowMask, every ray of the timed buffer"}
    t64, u64, v64 = sc.hits_f64(rays, o["tri_id"])
    h = hit & np.isfinite(t64)
    t_raw = o["t"][h].astype(np.float64) + np.float64(np.float32(1e-5))
    rel = np.abs(t_raw - t64[h]) / np.abs(t64[h])
    du = np.abs(o["u"][h].astype(np.float64) - u64[h])
    dv = np.abs(o["v"][h].astype(np.float64) - v64[h])
    err = {"hits": int(h.sum()),
           "t_max_rel": float(rel.max()) if h.any() else None, "t_p99_rel": float(np.percentile(rel, 99)) if h.any() else None,
           "u_max_abs": float(du.max()) if h.any() else None, "v_max_abs": float(dv.max()) if h.any() else None,
           "u_p99_abs": float(np.percentile(du, 99)) if h.any() else None,
           "v_p99_abs": float(np.percentile(dv, 99)) if h.any() else None,
           "oracle": "orc_hits_f64: double Moller-Trumbore on the hit triangle; t compared as stored t + 1e-5f"}
    return parity, err


def oracle_samples(groups, threads: int, target_s: float):
    """groups: per scene (oracle scene, [(rays, closest?, gpu result, gpu leaf ids or None), ...]), each
    ray set a sample of a batch item or request. Checks every sampled ray against the oracle (closest:
    triangle id, leaf id and the bits of t, u, v; any: the flag), then times the oracle's scalar
    traversal over the same samples for about target_s. -> (parity, Mrays/s, rays per pass)"""
    checked = bad = 0
    for sc, sets in groups:
        for rays, closest, got, nid in sets:
            checked += rays.size
            if closest:
                o = sc.trace_closest(rays, threads=threads)
                m = got["tri_id"] != o["tri_id"]
                if nid is not None:
                    m |= nid != o["nid"]
                for k in ("t", "u", "v"):
                    m |= got[k].view(np.uint32) != o[k].view(np.uint32)
                bad += int(m.sum())
            else:
                bad += int((got != sc.trace_any(rays, threads=threads)[0]).sum())
    # timing: one oracle call per scene over its concatenated samples
    merged = [(sc, np.concatenate([r for r, c, _, _ in sets if c]) if any(c for _, c, _, _ in sets) else None,
               np.concatenate([r for r, c, _, _ in sets if not c]) if any(not c for _, c, _, _ in sets) else None)
              for sc, sets in groups]
    per_pass = sum((a.size if a is not None else 0) + (b.size if b is not None else 0) for _, a, b in merged)
    passes = 0
    t0 = time.perf_counter()
    while True:
        for sc, a, b in merged:
            if a is not None:
                sc.trace_closest(a, threads=threads)
            if b is not None:
                sc.trace_any(b, threads=threads)
        passes += 1
        if time.perf_counter() - t0 >= target_s or passes >= 100:
            break
    dt = time.perf_counter() - t0
    return {"rays_checked": checked, "mismatches": bad}, passes * per_pass / dt / 1e6, per_pass, passes, dt


def cpu_config1(threads: int) -> dict:
    """The CPU method on the closed box (16 384 triangles): the builder restatement's wall time and the
    oracle's scalar closest-hit traversal of 256x256 jittered primary rays, on 1 thread and on `threads`."""
    import oracle
    from tim_raytracing_amd import host as H

    out = {}
    for th in sorted({1, threads}):
        t0 = time.perf_counter()
        b, f, w, h = H.workload(1, threads=th)
        build_s = time.perf_counter() - t0
        blob, ranges = b.fill_gpu_buffer()
        sc = oracle.Scene(blob, ranges, b.positions())
        rays = H.frustum_rays(f, w, h, seed=1)
        t0 = time.perf_counter()
        reps = 0
        while True:
            sc.trace_closest(rays, threads=th)
            reps += 1
            if time.perf_counter() - t0 > 1.0:
                break
        dt = time.perf_counter() - t0
        out[f"threads_{th}"] = {"build_s": round(build_s, 4), "trace_mrays_s": round(reps * rays.size / dt / 1e6, 3),
                                "rays": int(rays.size), "passes": reps}
    out["workload"] = "box-16K-primary-256x256 (16 384 tris, 256x256 jittered primary rays, t limit 100)"
    return out


def cpu_baseline(blob, ranges, pos, rays, threads: int, target_s: float, gpu=None):
    import oracle  # the CPU baseline leg is the only place bench.py touches the oracle

    sc = oracle.Scene(blob, ranges, pos)
    extra = {}
    if gpu is not None:
        extra["parity"], extra["fp64_error"] = check_against_oracle(sc, rays, *gpu, threads=threads)
    value, desc = timed_oracle(lambda r, th: sc.trace_closest(r, threads=th), rays, threads, target_s,
                               "primary rays, closest hit, bitstack restatement of traverseBvh")
    out = {"value": round(value, 4), "unit": "Mrays/s", "cores": threads, "kind": "port", "sample": desc,
           "cpu_model": cpu_model()}
    if threads > 1:
        one, one_desc = timed_oracle(lambda r, th: sc.trace_closest(r, threads=th), rays, 1, target_s / 5,
                                     "primary rays")
        out["single_thread_value"] = round(one, 4)
        out["single_thread_sample"] = one_desc
    out["config1_cpu_method"] = cpu_config1(threads)
    return out, extra


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; --gpus above 1 is not supported (rays "
                                   "are independent, so a multi-GPU run would split them into contiguous "
                                   "ranges over replicated scenes)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.mode == "batch":
        return bench_batch(args)
    if args.mode == "serve":
        return bench_serve(args)
    if args.mode == "incoherent":
        return bench_incoherent(args)
    if args.mode == "copy":
        return bench_copy(args)
    if args.mode == "tlas":
        return bench_tlas(args)

    from tim_raytracing_amd import bindings as B
    from tim_raytracing_amd import device as D
    from tim_raytracing_amd import host as H

    threads = host_threads(args.threads)
    t0 = time.perf_counter()
    builder, frustum, W, Hh = H.workload(2, scale=1.0, threads=threads)
    build_s = builder.stats().build_seconds
    blob, ranges = builder.fill_gpu_buffer()
    pos = builder.positions()
    rays_row = H.frustum_rays(frustum, W, Hh, seed=2)
    other_order = "row" if args.ray_order == "tile" else "tile"
    rays_tile = np.ascontiguousarray(rays_row[tile_order(W, Hh)])
    rays, rays_other = (rays_tile, rays_row) if args.ray_order == "tile" else (rays_row, rays_tile)
    n = rays.size
    setup_s = time.perf_counter() - t0

    dev = D.Device(0)
    dispatch_cfg = configure_dispatch(dev, args, "grid")
    scene = D.Scene(dev, blob, ranges, pos)
    info = scene.info()
    kname = kernel_name(args, int(info.max_depth))
    stream = dev.stream()
    sptr = stream.ptr
    d_rays = dev.upload(rays)
    d_hits = dev.alloc(16 * n)
    d_nid = dev.alloc(4 * n)
    d_masks = dev.alloc(4 * n)
    ev0, ev1 = dev.event(), dev.event()

    def closest():
        scene.dispatch(d_rays.ptr, n, B.RTQ_TRACE_CLOSEST, d_hits.ptr, d_nid.ptr, sptr)

    def timed(fn, warmup, steps):
        for _ in range(warmup):
            fn()
        stream.synchronize()
        ev0.record(sptr)
        for _ in range(steps):
            fn()
        ev1.record(sptr)
        return ev0.elapsed_ms(ev1) / steps

    # per-ray visit counts for the algorithmic bytes (counting variant, untimed)
    d_cnt = dev.alloc(12 * n)
    scene.dispatch_counted(d_rays.ptr, n, B.RTQ_TRACE_CLOSEST, d_hits.ptr, d_nid.ptr, d_cnt.ptr, sptr)
    stream.synchronize()
    cnt = d_cnt.read(np.uint32, 3 * n).astype(np.int64).reshape(n, 3)
    sum_inner, sum_leaf, sum_tri = (int(x) for x in cnt.sum(0))
    alg_bytes = 48 * n + 64 * sum_inner + 4 * sum_leaf + 48 * sum_tri
    d_cnt.free()

    ms = timed(closest, args.warmup, args.steps)
    hits = d_hits.read(B.HIT_DTYPE, n)  # the timed launch's records (its last step)
    nids = d_nid.read(np.uint32, n)
    hit_mask = hits["tri_id"] != B.RTQ_NO_HIT
    n_hit = int(hit_mask.sum())

    # shadow pass: one any-hit ray per hit toward the point light (light 0 of every leaf list)
    zero_sun = np.zeros(3, np.float32)
    shadow_ms = timed(lambda: scene.shadow_masks(d_rays.ptr, d_hits.ptr, d_nid.ptr, n, zero_sun, 0, d_masks.ptr, sptr),
                      max(1, args.warmup // 2), args.steps)
    masks = d_masks.read(np.uint32, n)

    # the camera pass form: rays made in the kernel from the frustum (no jitter), 8x8 pixel tiles per wave
    import ctypes as C
    camera_ms = timed(lambda: B.check_device("rtq_trace_camera", dev.lib.rtq_trace_camera(
        scene.handle, C.byref(frustum), W, Hh, 0, d_hits.ptr, d_nid.ptr, sptr)), max(1, args.warmup // 2), args.steps)

    # the full camera-pass record (cameraPass.comp TRACING_STEP): closest hit, shadow mask (sun + the hit
    # leaf's lights) and probe mask (8 any-hit segments) per pixel, 32 B per pixel
    sun = np.array([0.3, 0.2, -1.0], np.float32)
    sun /= np.linalg.norm(sun)
    smn, smx = builder.scene_box()
    lpf = B.Lpf((C.c_float * 3)(*[float(x) for x in smn]), (C.c_float * 3)(*[float(x) for x in smx]),
                (C.c_uint32 * 3)(16, 16, 16))
    d_records = dev.alloc(32 * n)
    record_ms = timed(lambda: B.check_device("rtq_trace_camera_record", dev.lib.rtq_trace_camera_record(
        scene.handle, C.byref(frustum), W, Hh, 0, sun.ctypes.data, B.RTQ_SHADOW_SUN, C.byref(lpf), d_records.ptr,
        sptr)), max(1, args.warmup // 2), args.steps)
    del d_records

    # the light-probe field's tracing step (genLightProbField.comp with TRACING_STEP): 64 rays per probe
    # of a 16^3 field over the scene box, each a closest hit plus its sun and leaf-light shadow rays
    probe_dirs = H.probe_directions(7)
    n_probe_rays = 16 * 16 * 16 * B.RTQ_PROBE_RAYS
    d_probe_records = dev.alloc(32 * n_probe_rays)
    probe_ms = timed(lambda: B.check_device("rtq_trace_probe_field", dev.lib.rtq_trace_probe_field(
        scene.handle, C.byref(lpf), probe_dirs.ctypes.data, sun.ctypes.data, B.RTQ_SHADOW_SUN, d_probe_records.ptr,
        sptr)), max(1, args.warmup // 2), args.steps)
    del d_probe_records

    # the same rays in the other buffer order, for comparison (same kernel, same per-ray work)
    d_rays.write(rays_other)
    other_ms = timed(closest, max(1, args.warmup // 2), args.steps)

    alg_gbs = alg_bytes / (ms * 1e-3) / 1e9
    algorithmic = {
        "bytes_per_launch": int(alg_bytes),
        "achieved_gb_s": round(alg_gbs, 1),
        "vs_hbm_peak": round(alg_gbs / HBM_PEAK_GBS, 4),
        "mean_bytes_per_ray": round(alg_bytes / n, 1),
        "mean_inner_nodes": round(sum_inner / n, 2),
        "mean_leaves": round(sum_leaf / n, 3),
        "mean_triangle_tests": round(sum_tri / n, 2),
        "hbm_roofline_mrays_s": round(HBM_PEAK_GBS * 1e9 / (alg_bytes / n) / 1e6, 1),
        # reported, not a roof: about 50 flops per inner node (two slab tests) and 45 per triangle test
        "flops_per_ray": round((50.0 * sum_inner + 45.0 * sum_tri) / n, 1),
        "note": "B = 32 + 16 + 64 N_inner + 4 N_leaf + 48 N_tri per ray at the reference layout, counts from the "
                "counting kernel (tests check them against the oracle); the 119 MB scene stays in the 256 MB "
                "Infinity Cache, so HBM sees about 1 % of these bytes and they are not a fraction of the HBM roof",
    }
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
            "ray_order": args.ray_order,
            "bvh_max_depth": int(info.max_depth),
            "scene_device_bytes": int(info.device_bytes),
        },
        "roofline": dict(roofline("spheres-1M-primary-1920x1080", f"{kname}@{args.ray_order}", ms, algorithmic),
                         dispatch=dispatch_cfg),
        "hit_rate": round(n_hit / n, 4),
        "shadow": {
            "rays": n_hit,
            "ms_per_step": round(shadow_ms, 4),
            "anyhit_mrays_s": round(n_hit / (shadow_ms * 1e-3) / 1e6, 2),
            "occluded_frac": round(float(((masks >> 1) & 1)[hit_mask].mean()) if n_hit else 0.0, 4),
        },
        "camera_pass": {"kernel": "kTraceCamera", "ms_per_step": round(camera_ms, 4),
                        "mrays_s": round(n / (camera_ms * 1e-3) / 1e6, 2)},
        "record_pass": {"kernel": "kTraceRecord", "ms_per_step": round(record_ms, 4),
                        "mpixels_s": round(n / (record_ms * 1e-3) / 1e6, 2),
                        "per_pixel": "closest hit + sun and leaf-light shadow rays + 8 probe segments (16^3 field)"},
        "probe_pass": {"kernel": "kTraceProbes", "ms_per_step": round(probe_ms, 4), "rays": n_probe_rays,
                       "mrays_s": round(n_probe_rays / (probe_ms * 1e-3) / 1e6, 2),
                       "per_ray": "closest hit + sun and leaf-light shadow rays, 64 rays per probe of a 16^3 field"},
        "other_ray_order": {"ray_order": other_order, "ms_per_step": round(other_ms, 4),
                            "mrays_s": round(n / (other_ms * 1e-3) / 1e6, 2)},
        "host": {"threads": threads, "bvh_build_s": round(build_s, 2), "setup_s": round(setup_s, 2)},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"], checks = cpu_baseline(blob, ranges, pos, rays, threads, args.cpu_seconds,
                                                      gpu=(hits, nids, masks))
        result.update(checks)
    print(json.dumps(result), flush=True)


def bench_copy(args):
    """1 GiB device-to-device copy with the library's streaming kernel (kCopy16): the known-size
    reference that the FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    from tim_raytracing_amd import bindings as B
    from tim_raytracing_amd import device as D

    dev = D.Device(0)
    nbytes = 1 << 30
    src, dst = dev.alloc(nbytes), dev.alloc(nbytes)
    stream = dev.stream()
    ev0, ev1 = dev.event(), dev.event()

    def copy():
        B.check_device("rtq_device_copy", dev.lib.rtq_device_copy(dev.handle, dst.ptr, src.ptr, nbytes, stream.ptr))

    for _ in range(args.warmup):
        copy()
    stream.synchronize()
    ev0.record(stream.ptr)
    for _ in range(args.steps):
        copy()
    ev1.record(stream.ptr)
    ms = ev0.elapsed_ms(ev1) / args.steps
    # link rates with page-locked host memory (what served requests move per ray)
    hbytes = 256 << 20
    host = dev.host_alloc(hbytes)
    host.view(np.uint8)[:] = 1

    def link(direction):
        fn = (lambda: B.check_device("rtq_copy_to_device", dev.lib.rtq_copy_to_device(
            dev.handle, src.ptr, host.ptr, hbytes, stream.ptr))) if direction == "h2d" else \
            (lambda: B.check_device("rtq_read_hits", dev.lib.rtq_read_hits(dev.handle, src.ptr, host.ptr, hbytes,
                                                                           stream.ptr)))
        fn()
        stream.synchronize()
        ev0.record(stream.ptr)
        for _ in range(5):
            fn()
        ev1.record(stream.ptr)
        return round(5 * hbytes / (ev0.elapsed_ms(ev1) * 1e-3) / 1e9, 1)

    h2d, d2h = link("h2d"), link("d2h")
    # both directions at once, each from its own host thread on its own stream (what the served form does
    # with rays in and results out); ctypes releases the interpreter lock during the copies
    import threading

    host2 = dev.host_alloc(hbytes)
    stream2 = dev.stream()

    def loop(direction, reps):
        for _ in range(reps):
            if direction == "h2d":
                B.check_device("rtq_copy_to_device", dev.lib.rtq_copy_to_device(dev.handle, src.ptr, host.ptr,
                                                                                  hbytes, stream.ptr))
            else:
                B.check_device("rtq_read_hits", dev.lib.rtq_read_hits(dev.handle, dst.ptr, host2.ptr, hbytes,
                                                                        stream2.ptr))

    def both(reps):
        th = [threading.Thread(target=loop, args=(d, reps)) for d in ("h2d", "d2h")]
        for t in th:
            t.start()
        for t in th:
            t.join()

    both(1)
    t0 = time.perf_counter()
    both(5)
    duplex = round(5 * hbytes / (time.perf_counter() - t0) / 1e9, 1)
    host.free()
    host2.free()
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": round(2 * nbytes / (ms * 1e-3) / 1e9, 1),
                      "unit": "GB/s", "bytes_copied": nbytes, "ms_per_copy": round(ms, 4), "kernel": "kCopy16",
                      "steps": args.steps, "warmup": args.warmup,
                      "pinned_link_gb_s": {"h2d": h2d, "d2h": d2h, "each_way_both_at_once": duplex,
                                           "bytes": hbytes}}), flush=True)


def bench_tlas(args):
    """Two-level any hit: 1024 BLAS instances of 1000 small triangles each (clusters in [-100, 100]^3) and
    200 top-level triangles, 4 Mi rays (from afar toward the clusters, from inside BLAS boxes, onto the
    top-level triangles, random), one rtq_dispatch_rays(RTQ_TRACE_ANY) per step."""
    from tim_raytracing_amd import bindings as B
    from tim_raytracing_amd import device as D
    from tim_raytracing_amd import host as H

    threads = host_threads(args.threads)
    t0 = time.perf_counter()
    b = H.two_level_scene(11, n_blas=1024, tris_per_blas=1000, own_triangles=200, spread=100.0)
    b.build(threads)
    build_s = time.perf_counter() - t0
    dev = D.Device(0)
    # lane refill without a sort was the fastest of the dispatch modes on this workload
    dispatch_cfg = configure_dispatch(dev, args, "persistent", auto_sort=0)
    blob, ranges = b.fill_gpu_buffer()
    pos = b.positions()
    scene = D.Scene(dev, blob, ranges, pos)
    n = 4 << 20
    rays = H.two_level_rays(b, 12, n).view(B.RAY_DTYPE).reshape(-1)
    d_r, d_o = dev.upload(rays), dev.alloc(4 * n)
    stream = dev.stream()
    sptr = stream.ptr
    ev0, ev1 = dev.event(), dev.event()

    def step():
        scene.dispatch(d_r.ptr, n, B.RTQ_TRACE_ANY, d_o.ptr, None, sptr)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    ev0.record(sptr)
    for _ in range(args.steps):
        step()
    ev1.record(sptr)
    ms = ev0.elapsed_ms(ev1) / args.steps
    got = d_o.read(np.uint32, n)
    depth = int(scene.info().max_depth)
    kname = kernel_name(args, depth, any_hit=True)
    result = {
        "metric": "two-level any-hit Mrays/s", "value": round(n / (ms * 1e-3) / 1e6, 2), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "tlas-1024-blas-any", "blas_instances": 1024, "triangles_per_blas": 1000,
                   "top_level_triangles": 200, "rays": n, "depth": depth},
        "occluded_frac": round(float(got.mean()), 4),
        "roofline": roofline("tlas-1024-blas-any", kname, ms),
        "dispatch": dispatch_cfg, "kernel": kname, "host": {"threads": threads, "build_s": round(build_s, 1)},
    }
    if not args.no_cpu_baseline:
        import oracle  # the CPU baseline leg is the only place bench.py touches the oracle

        sc = oracle.Scene(blob, ranges, pos)
        want = sc.trace_any_tlas(rays, threads=threads)
        result["parity"] = {"rays_checked": int(n), "mismatches": int((got != want).sum()),
                            "checker": "oracle traverseTlasFast / tlas_collide_fast restatement, every ray"}
        value, desc = timed_oracle(lambda r, th: sc.trace_any_tlas(r, threads=th), rays, threads, args.cpu_seconds,
                                   "two-level any hit, bitstack restatement of traverseTlasFast")
        result["cpu_baseline"] = {"value": round(value, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
                                  "sample": desc, "cpu_model": cpu_model()}
    print(json.dumps(result), flush=True)


def bench_batch(args):
    """Batched form: 256 scenes of LogU[1K, 2M] triangles, LogU[4K, 1M] rays each (half primary,
    half bounce), one persistent launch (autotuned, hipGraph replay) vs one launch per item."""
    from tim_raytracing_amd import bindings as B
    from tim_raytracing_amd import device as D

    from tim_raytracing_amd import host as H

    threads = host_threads(args.threads)
    slots = 256
    builders, ray_counts = H.batch_workload(args.scale, slots)
    distinct = args.distinct or slots
    t0 = time.perf_counter()
    pool = H.build_many(builders[:distinct], threads, args.cache_dir, progress=lambda m: print(m, file=sys.stderr,
                                                                                               flush=True))
    build_s = time.perf_counter() - t0
    dev = D.Device(0)
    scenes = [D.Scene(dev, blob, r, pos) for blob, r, pos in pool]
    items, keep = [], []
    total = 0
    for i in range(slots):
        si = i % distinct
        n = int(ray_counts[i])
        rays = H.batch_item_rays(builders[si], i, n)
        d_r = dev.upload(rays)
        d_o = dev.alloc(16 * n)
        d_n = dev.alloc(4 * n)
        keep.append((d_r, d_o, d_n))
        items.append((si, B.RTQ_TRACE_CLOSEST, d_r.ptr, n, d_o.ptr, d_n.ptr))
        total += n
    batch = D.Batch(dev, scenes, items)
    stream = dev.stream()
    sptr = stream.ptr
    ev0, ev1 = dev.event(), dev.event()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        stream.synchronize()
        ev0.record(sptr)
        for _ in range(args.steps):
            fn()
        ev1.record(sptr)
        return ev0.elapsed_ms(ev1) / args.steps

    plain_ms = timed(lambda: batch.launch_plain(sptr))
    if args.batch_params:
        ch, bl, sd, lw = (int(x) for x in args.batch_params.split(","))
        batch.set_params(chunk_rays=ch, blocks_per_cu=bl, stack_depth=sd, leaf_wait=lw)
        tuned_ms = 0.0
    else:
        print(f"plain: {plain_ms:.3f} ms per batch; autotuning", file=sys.stderr, flush=True)
        best, tuned_ms = batch.autotune(reps=2, stream=sptr)
    batch.set_params(use_graph=1)
    batched_ms = timed(lambda: batch.launch(sptr))
    p = batch.params()
    kname = f"{'kBatchRefill' if p.chunk_rays == B.RTQ_BATCH_REFILL else 'kBatchPersistent'}<{p.stack_depth}>"
    result = {
        "metric": "batched closest-hit Mrays/s vs one launch per item",
        "value": round(total / (batched_ms * 1e-3) / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(batched_ms, 4),
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch-256-scenes", "scene_slots": slots, "distinct_bvhs": distinct,
                   "triangles_per_slot": "LogU[1e3, %g]" % (2e6 * args.scale),
                   "distinct_triangles": int(sum(r.triangles.size // 12 for _, r, _ in pool)),
                   "rays": int(total), "rays_per_item": "LogU[4096, 1048576]", "mix": "50% primary, 50% bounce"},
        "plain_mrays_s": round(total / (plain_ms * 1e-3) / 1e6, 2), "plain_ms": round(plain_ms, 4),
        "speedup_vs_plain": round(plain_ms / batched_ms, 3),
        "autotuned": {"chunk_rays": p.chunk_rays, "blocks_per_cu": p.blocks_per_cu, "stack_depth": p.stack_depth,
                      "use_graph": p.use_graph, "leaf_wait": p.leaf_wait, "tuning_ms": round(tuned_ms, 4)},
        "roofline": roofline("batch-256-scenes" if distinct == slots else f"batch-256-scenes-{distinct}-bvhs", kname,
                             batched_ms),
        "host": {"threads": threads, "build_s": round(build_s, 1)},
    }
    if not args.no_cpu_baseline:
        import oracle  # the CPU baseline leg is the only place bench.py touches the oracle

        # the graph-replayed launch's results of every item, sampled (4096 strided rays per item)
        groups = {}
        for i in range(slots):
            si = i % distinct
            d_r, d_o, d_n = keep[i]
            n = int(ray_counts[i])
            idx = np.arange(0, n, max(1, n // 4096))[:4096]
            rays = H.batch_item_rays(builders[si], i, n)[idx]
            hits = d_o.read(B.HIT_DTYPE, n)[idx]
            nid = d_n.read(np.uint32, n)[idx]
            groups.setdefault(si, []).append((np.ascontiguousarray(rays), True, hits, nid))
        scs = [(oracle.Scene(*pool[si]), sets) for si, sets in groups.items()]
        parity, rate, per_pass, passes, dt = oracle_samples(scs, threads, args.cpu_seconds)
        result["parity"] = dict(parity, checker="oracle traverseBvh on 4096 strided rays of every item")
        result["cpu_baseline"] = {"value": round(rate, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
                                  "sample": f"{per_pass} rays (4096 strided rays of each of the {slots} items) x "
                                            f"{passes} passes, closest hit, bitstack restatement of traverseBvh, "
                                            f"{threads} threads, {dt:.1f} s", "cpu_model": cpu_model()}
    print(json.dumps(result), flush=True)


def bench_serve(args):
    """Served form: 64 resident scenes of LogU[10K, 3M] triangles, a Poisson stream of requests of
    LogU[1K, 2M] rays (half closest, half any hit) at 1.5x the plain capacity; continuous batching vs
    one launch per request under the same HBM budget. Request rays and results live in page-locked
    host memory, so the link moves them by DMA. Both forms are also run saturated (every request
    submitted at once) to compare their capacities."""
    import threading

    from tim_raytracing_amd import bindings as B
    from tim_raytracing_amd import device as D

    from tim_raytracing_amd import host as H

    threads = host_threads(args.threads)
    slots = 64
    nreq = args.requests
    builders, rng, req_sizes, req_scene, req_any = H.serve_workload(args.scale, slots, nreq)
    distinct = args.distinct or slots
    t0 = time.perf_counter()
    pool = H.build_many(builders[:distinct], threads, args.cache_dir, progress=lambda m: print(m, file=sys.stderr,
                                                                                               flush=True))
    build_s = time.perf_counter() - t0
    dev = D.Device(0)
    # every slot is its own resident copy in HBM, also where slots share a distinct BVH
    scenes = [D.Scene(dev, *pool[i % distinct]) for i in range(slots)]
    max_depth = max(int(s.info().max_depth) for s in scenes)
    pool_rays = 1 << 21
    pinned_pools = []
    for i in range(distinct):
        hb = dev.host_alloc(32 * pool_rays)
        hb.view(B.RAY_DTYPE, pool_rays)[:] = builders[i].surface_rays(seed=50 + i, count=pool_rays)
        pinned_pools.append(hb)
    pools = [hb.view(B.RAY_DTYPE, pool_rays) for hb in pinned_pools]
    req_off = [int(rng.integers(0, pool_rays - n + 1)) for n in req_sizes]
    out_off = np.concatenate([[0], np.cumsum(np.where(req_any, 4, 16) * req_sizes)]).astype(np.int64)
    out_arena = dev.host_alloc(int(out_off[-1]))
    budget = 2 << 30
    params = B.BatchParams(args.chunk_rays, args.blocks_per_cu or 8, stack_class(max_depth), 0, args.leaf_wait)

    def run(plain, arrivals):
        srv = D.Server(dev, scenes, budget_bytes=budget, plain=plain, max_batch_rays=args.max_batch_rays,
                       params=params, max_inflight=args.max_inflight)
        handles = [None] * nreq
        start = time.perf_counter()

        def client():
            for k in range(nreq):
                if arrivals is not None:
                    wait = start + arrivals[k] - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                si = int(req_scene[k])
                n = int(req_sizes[k])
                rays = pools[si % distinct][req_off[k]:req_off[k] + n]
                if req_any[k]:
                    out = out_arena.view(np.uint32, n, int(out_off[k]))
                    handles[k] = srv.submit(si, B.RTQ_TRACE_ANY, rays, out=out)
                else:
                    out = out_arena.view(B.HIT_DTYPE, n, int(out_off[k]))
                    handles[k] = srv.submit(si, B.RTQ_TRACE_CLOSEST, rays, out=out, with_nid=False)

        th = threading.Thread(target=client)
        th.start()
        th.join()
        lat = []
        for h in handles:
            assert h.wait() == B.RTQ_OK
            lat.append(h.stats().latency_ms)
        end = time.perf_counter()
        batches = len({h.stats().batch_id for h in handles})
        for h in handles:
            h.release()
        srv.close()
        lat = np.array(lat)
        return {"mrays_s": round(float(req_sizes.sum()) / (end - start) / 1e6, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "makespan_s": round(end - start, 3), "launches": batches}

    # order: warm-up (served), plain and served capacities (everything submitted at once), then plain and
    # served at 1.5x the plain capacity; the served runs are the 1st, 3rd and 5th groups of batch launches
    run(False, None)  # warm-up: first launches of every kernel and scene
    plain_cap = run(True, None)
    served_cap = run(False, None)
    rate = 1.5 * plain_cap["mrays_s"] * 1e6 / float(req_sizes.mean())
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    plain = run(True, arrivals)
    served = run(False, arrivals)

    # The same requests with rays and results in device memory (rays made and results consumed by other
    # GPU passes): rtq_server_submit_device traces them in place, so nothing crosses the host link and
    # both forms are bound by their kernels and launches. Same order of runs; the at-load served
    # results must equal the host-buffer run's on every ray.
    d_pools = [dev.upload(pools[i]) for i in range(distinct)]
    out_bytes = np.where(req_any, 4, 16) * req_sizes
    dout_off = np.concatenate([[0], np.cumsum((out_bytes + 255) // 256 * 256)]).astype(np.int64)  # 256 B aligned
    d_arena = dev.alloc(int(dout_off[-1]))

    def run_dev(plain_form, arrivals_dev):
        srv = D.Server(dev, scenes, budget_bytes=budget, plain=plain_form, max_batch_rays=args.max_batch_rays,
                       params=params, max_inflight=args.max_inflight)
        handles = [None] * nreq
        start = time.perf_counter()
        for k in range(nreq):
            if arrivals_dev is not None:
                wait = start + arrivals_dev[k] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
            si = int(req_scene[k])
            n = int(req_sizes[k])
            rays = D.DeviceView(d_pools[si % distinct], 32 * req_off[k], 32 * n)
            flags = B.RTQ_TRACE_ANY if req_any[k] else B.RTQ_TRACE_CLOSEST
            out = D.DeviceView(d_arena, int(dout_off[k]), int(out_bytes[k]))
            handles[k] = srv.submit_device(si, flags, rays, n, out)
        lat = []
        for h in handles:
            assert h.wait() == B.RTQ_OK
            lat.append(h.stats().latency_ms)
        end = time.perf_counter()
        batches = len({h.stats().batch_id for h in handles})
        for h in handles:
            h.release()
        srv.close()
        lat = np.array(lat)
        return {"mrays_s": round(float(req_sizes.sum()) / (end - start) / 1e6, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "makespan_s": round(end - start, 3), "launches": batches}

    run_dev(False, None)
    dplain_cap = run_dev(True, None)
    dserved_cap = run_dev(False, None)
    drate = 1.5 * dplain_cap["mrays_s"] * 1e6 / float(req_sizes.mean())
    darrivals = np.cumsum(np.random.default_rng(7).exponential(1.0 / drate, nreq))
    dplain = run_dev(True, darrivals)
    dserved = run_dev(False, darrivals)
    dres = d_arena.read(np.uint8, int(dout_off[-1]))
    hres = out_arena.view(np.uint8, int(out_off[-1]))
    same = all(np.array_equal(dres[dout_off[k]:dout_off[k] + out_bytes[k]], hres[out_off[k]:out_off[k] + out_bytes[k]])
               for k in range(nreq))
    del dres
    device_requests = {
        "served": dserved, "plain": dplain,
        "speedup_vs_plain": round(dserved["mrays_s"] / dplain["mrays_s"], 3) if dplain["mrays_s"] else None,
        "capacity": {"served": dserved_cap, "plain": dplain_cap,
                     "speedup_vs_plain": round(dserved_cap["mrays_s"] / dplain_cap["mrays_s"], 3)},
        "arrival_rate_per_s": round(drate, 2),
        "results_equal_host_requests": same,
        "requests": "the same 1000 requests, rays and results in HBM (rtq_server_submit_device); Poisson arrivals "
                    "at 1.5x the plain form's capacity on such requests",
    }
    del d_pools, d_arena
    checks = {}
    if not args.no_cpu_baseline:
        import oracle  # the CPU baseline leg is the only place bench.py touches the oracle

        # the at-load served results of every request, sampled (4096 strided rays per request)
        groups = {}
        for k in range(nreq):
            si = int(req_scene[k])
            n = int(req_sizes[k])
            idx = np.arange(0, n, max(1, n // 4096))[:4096]
            rays = np.ascontiguousarray(pools[si % distinct][req_off[k]:req_off[k] + n][idx])
            if req_any[k]:
                got = out_arena.view(np.uint32, n, int(out_off[k]))[idx].copy()
                groups.setdefault(si % distinct, []).append((rays, False, got, None))
            else:
                got = out_arena.view(B.HIT_DTYPE, n, int(out_off[k]))[idx].copy()
                groups.setdefault(si % distinct, []).append((rays, True, got, None))
        scs = [(oracle.Scene(*pool[si]), sets) for si, sets in groups.items()]
        parity, rate, per_pass, passes, dt = oracle_samples(scs, threads, args.cpu_seconds)
        checks["parity"] = dict(parity, checker="oracle traverseBvh / traverseBvhFast on 4096 strided rays of "
                                                "every served request (at-load run)")
        checks["cpu_baseline"] = {"value": round(rate, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
                                  "sample": f"{per_pass} rays (4096 strided rays of each of the {nreq} requests, "
                                            f"half closest, half any hit) x {passes} passes, bitstack restatement "
                                            f"of traverseBvh / traverseBvhFast, {threads} threads, {dt:.1f} s",
                                  "cpu_model": cpu_model()}
    kname = f"{'kBatchRefill' if params.chunk_rays == B.RTQ_BATCH_REFILL else 'kBatchPersistent'}<{params.stack_depth}>"
    result = {
        "metric": "served aggregate Mrays/s and p50/p99 latency vs one launch per request",
        "value": served["mrays_s"], "unit": "Mrays/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "serve-64-scenes-poisson", "scene_slots": slots, "distinct_bvhs": distinct,
                   "triangles_per_slot": "LogU[1e4, %g]" % (3e6 * args.scale), "requests": nreq,
                   "rays": int(req_sizes.sum()), "rays_per_request": "LogU[1000, 2e6]",
                   "mix": "50% closest, 50% any", "arrival_rate_per_s": round(rate, 2),
                   "offered_load": "1.5x plain capacity", "hbm_budget_bytes": budget,
                   "host_buffers": "page-locked", "max_batch_rays": args.max_batch_rays,
                   "max_inflight": args.max_inflight or 3,
                   "params": {"chunk_rays": params.chunk_rays, "blocks_per_cu": params.blocks_per_cu,
                              "stack_depth": params.stack_depth}},
        "served": served, "plain": plain,
        "speedup_vs_plain": round(served["mrays_s"] / plain["mrays_s"], 3) if plain["mrays_s"] else None,
        "capacity": {"served": served_cap, "plain": plain_cap,
                     "speedup_vs_plain": round(served_cap["mrays_s"] / plain_cap["mrays_s"], 3)},
        "device_requests": device_requests,
        # over the saturated served run's makespan: the VALU instructions of all its launches (committed
        # counters of one such run) per second of wall time, a lower bound of the kernels' own issue rate
        "roofline": roofline("serve-64-scenes-poisson", f"{kname}@saturated-stream",
                             served_cap["makespan_s"] * 1e3),
        "host": {"threads": threads, "build_s": round(build_s, 1)},
    }
    result.update(checks)
    print(json.dumps(result), flush=True)


def bench_incoherent(args):
    """3 Mi-triangle Gaussian-mixture soup with 4096 point lights: 4 Mi diffuse bounce rays from
    surfaces, a second bounce from their hits, and 16x16 tile culling of the lights at 1920x1080."""
    from tim_raytracing_amd import bindings as B
    from tim_raytracing_amd import device as D
    from tim_raytracing_amd import host as H

    threads = host_threads(args.threads)
    t0 = time.perf_counter()
    b, f, W, Hh = H.workload(3, scale=args.scale, threads=threads)
    build_s = time.perf_counter() - t0
    dev = D.Device(0)
    dispatch_cfg = configure_dispatch(dev, args, "persistent", auto_sort=1)
    scene = D.Scene.from_builder(dev, b)
    n = 4 << 20
    rays1 = b.surface_rays(seed=3, count=n)
    stream = dev.stream()
    sptr = stream.ptr
    ev0, ev1 = dev.event(), dev.event()
    d_h = dev.alloc(16 * n)
    d_n = dev.alloc(4 * n)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        stream.synchronize()
        ev0.record(sptr)
        for _ in range(args.steps):
            fn()
        ev1.record(sptr)
        return ev0.elapsed_ms(ev1) / args.steps

    def trace(rays):
        d_r = dev.upload(rays)
        ms = timed(lambda: scene.dispatch(d_r.ptr, n, B.RTQ_TRACE_CLOSEST, d_h.ptr, d_n.ptr, sptr))
        return ms, d_h.read(B.HIT_DTYPE, n), d_n.read(np.uint32, n)

    ms1, hits1, nid1 = trace(rays1)
    rays2 = b.bounce_rays(4, rays1, hits1)
    ms2, hits2, nid2 = trace(rays2)
    ntiles = ((W + 15) // 16) * ((Hh + 15) // 16)
    offs, lists = scene.cull_lights(f, W, Hh)
    # culling on the stream alone (asynchronous entry point, no host synchronisation in the timed loop)
    d_off, d_lists = dev.alloc(4 * (ntiles + 1)), dev.alloc(4 * max(1, lists.size))
    cull_ms = timed(lambda: scene.cull_lights_async(f, W, Hh, d_off.ptr, d_lists.ptr, int(lists.size), sptr))
    g_off, g_lists = d_off.read(np.uint32, ntiles + 1), d_lists.read(np.uint32, max(1, lists.size))[:lists.size]
    assert np.array_equal(g_off, offs) and np.array_equal(g_lists, lists)
    kname = kernel_name(args, int(scene.info().max_depth))
    workload = "soup-3M-bounce2-cull4096"
    n_lights = int(scene.info().num_lights)
    result = {
        "metric": "incoherent closest-hit Mrays/s (2 diffuse bounces) + tile light culling",
        "value": round(2 * n / ((ms1 + ms2) * 1e-3) / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms1 + ms2, 4), "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "triangles": int(b.stats().num_triangles),
                   "lights": n_lights, "rays_per_bounce": n, "tiles": ntiles},
        "bounce1_mrays_s": round(n / (ms1 * 1e-3) / 1e6, 2), "bounce2_mrays_s": round(n / (ms2 * 1e-3) / 1e6, 2),
        "bounce1_hit_rate": round(float((hits1["tri_id"] != B.RTQ_NO_HIT).mean()), 4),
        "bounce2_hit_rate": round(float((hits2["tri_id"] != B.RTQ_NO_HIT).mean()), 4),
        "roofline": {"bounce1": roofline(workload, f"{kname}@bounce1", ms1),
                     "bounce2": roofline(workload, f"{kname}@bounce2", ms2)},
        "cull": {"ms_per_call": round(cull_ms, 4), "tile_light_pairs": int(lists.size),
                 "tile_light_tests_per_s": round(ntiles * n_lights / (cull_ms * 1e-3), 1),
                 "kernels": "kCullMask + kScan + kCullEmit, stream-ordered scratch, no host synchronisation",
                 "roofline": roofline(workload, "cull", cull_ms)},
        "dispatch": dispatch_cfg, "kernel": kname,
        "host": {"threads": threads, "build_s": round(build_s, 1)},
    }
    if not args.no_cpu_baseline:
        import oracle  # the CPU baseline leg is the only place bench.py touches the oracle

        blob, ranges = b.fill_gpu_buffer()
        sc = oracle.Scene(blob, ranges, b.positions())
        parity = {}
        for name, rays, hits, nid in (("bounce1", rays1, hits1, nid1), ("bounce2", rays2, hits2, nid2)):
            o = sc.trace_closest(rays, threads=threads)
            bad = (hits["tri_id"] != o["tri_id"]) | (nid != o["nid"])
            for k in ("t", "u", "v"):
                bad |= hits[k].view(np.uint32) != o[k].view(np.uint32)
            parity[name] = {"rays_checked": int(rays.size), "mismatches": int(bad.sum())}
        t0 = time.perf_counter()
        o_off, o_lists = sc.cull_lights(f, W, Hh)
        cull_cpu_s = time.perf_counter() - t0
        parity["cull"] = {"tiles_checked": ntiles,
                          "mismatches": int((not np.array_equal(o_off, offs)) or (not np.array_equal(o_lists, lists)))}
        result["parity"] = parity
        value, desc = timed_oracle(lambda r, th: sc.trace_closest(r, threads=th), rays1, threads, args.cpu_seconds,
                                   "bounce-1 diffuse rays, closest hit, bitstack restatement of traverseBvh")
        result["cpu_baseline"] = {"value": round(value, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
                                  "sample": desc, "cpu_model": cpu_model(),
                                  "cull": {"ms": round(cull_cpu_s * 1e3, 2), "threads": 1,
                                           "sample": f"all {ntiles} tiles x {n_lights} lights, orc_cull_lights"}}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
