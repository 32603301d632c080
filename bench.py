"""Throughput benchmark of the GPU path tracer.

One step = one full-frame render of the workload at its full spp (every
pixel, every sample, up to max_bounces scatter vertices), inputs already
resident in HBM.  The default workload is `surface1m` (1024x1024, 256 spp,
8 bounces, 1.31M-triangle mesh + spheres + rectangles, SAH BVH); pass
--workload volume128 / cloud512 for the volume scenes.

Prints one JSON line: samples/s (value), rays/s, the dominant kernel's
HBM roofline fraction (algorithmic bytes from a separate counting run divided
by its HIP-event-timed duration) and the CPU baseline (the oracle timed on this
host's cores on a bounded crop of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per counted event (units defined in ne_counters)
BYTES = {"density_lookups": 32, "majorant_cells": 4, "bvh_nodes": 64, "prim_tests": 48}


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def cpu_baseline(wl, budget_s: float):
    """Oracle (CPU restatement, FP32, same counter RNG and estimator) on a
    centred crop sized for about budget_s seconds, samples/s."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    w = O.World(wl.scene)
    build_s = time.perf_counter() - t0
    cx, cy = wl.width // 2, wl.height // 2

    def run(side):
        rect = (cx - side // 2, cy - side // 2, cx - side // 2 + side, cy - side // 2 + side)
        t = time.perf_counter()
        w.render(wl.camera, wl.width, wl.height, wl.spp, wl.max_bounces, rect=rect, seed=1, threads=threads)
        return time.perf_counter() - t, side * side * wl.spp

    dt, n = run(8)
    rate = n / max(dt, 1e-9)
    side = int(max(8, min(256, (budget_s * rate / wl.spp) ** 0.5)))
    side = max(8, side - side % 8)
    dt, n = run(side)
    return {"value": n / dt, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"{side}x{side} centre crop of the {wl.width}x{wl.height} frame at {wl.spp} spp, "
                      f"{wl.max_bounces} bounces ({n} paths, {dt:.1f} s; oracle scene build {build_s:.1f} s "
                      f"not timed)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="surface1m", choices=["surface1m", "volume128", "cloud512"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--samples-per-thread", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the path shards by "
              "pixel tiles / sample ranges with a final FP64 reduce, which is not part of this build", file=sys.stderr)
        sys.exit(2)

    from narvalengine_amd import Context, configs
    t0 = time.perf_counter()
    wl = configs.CONFIGS[args.workload]()
    log(f"workload {wl.name}: {wl.width}x{wl.height} spp {wl.spp} bounces {wl.max_bounces} "
        f"(scene built in {time.perf_counter() - t0:.1f} s)")
    ctx = Context(0)
    g = ctx.upload(wl.scene)
    info = g.info()
    log(f"scene on device: {info}")
    film = ctx.film(wl.width, wl.height)
    kw = dict(spp=wl.spp, max_bounces=wl.max_bounces, seed=1, samples_per_thread=args.samples_per_thread,
              block_size=args.block)
    for i in range(args.warmup):
        ctx.render(g, wl.camera, film, sample_offset=(1000 + i) * wl.spp, **kw)
        log(f"warmup {i + 1}/{args.warmup} done")
    ctx.synchronize()
    kernel_ms = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        ctx.render(g, wl.camera, film, sample_offset=k * wl.spp, sync=False, timed=True, **kw)
        kernel_ms.append(ctx.kernel_ms())
        log(f"step {k + 1}/{args.steps}: kernel {kernel_ms[-1]:.1f} ms")
    ctx.synchronize()
    wall = time.perf_counter() - t_start
    paths_per_step = wl.width * wl.height * wl.spp
    value = args.steps * paths_per_step / wall
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3

    roofline = None
    rays = None
    counters = None
    if not args.no_count:
        cfilm = ctx.film(wl.width, wl.height)
        ctx.render(g, wl.camera, cfilm, count=True, **kw)
        counters = ctx.counters()
        log(f"counters: {counters}")
        rays = counters["ext_rays"] + counters["shadow_rays"]
        film_bytes = 48 * wl.width * wl.height  # FP64 RGB read + write per pixel
        algo = sum(counters[k] * b for k, b in BYTES.items()) + film_bytes
        achieved = algo / avg_kernel_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                    "kernel": "k_render", "algorithmic_bytes_per_launch": int(algo),
                    "avg_launch_ms": round(avg_kernel_s * 1e3, 3)}
    cpu = None if args.no_cpu else cpu_baseline(wl, args.cpu_seconds)
    out = {
        "metric": "samples/sec", "value": round(value, 1), "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": wl.name, "width": wl.width, "height": wl.height, "spp": wl.spp,
                   "max_bounces": wl.max_bounces, "paths_per_step": paths_per_step, "accumulation": "f64"},
        "rays_per_sec": round(rays / avg_kernel_s, 1) if rays else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "counters_per_step": counters,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
