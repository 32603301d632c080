"""Throughput benchmark of the GPU path tracer.

One step = one full-frame render of the workload at its full spp (every
pixel, every sample, up to max_bounces scatter vertices), inputs already
resident in HBM.  The default workload is `cloud512` (512^3 cumulus grid with
16^3 majorant cells, 1920x1080, 256 spp, up to 128 bounces, ratio-tracked
shadow rays); pass --workload surface1m / volume128 for the other image
scenes, or batch48 for the 48-job batch in one persistent graph launch (one
step = one launch over every job at full spp).

Prints one JSON line: samples/s (value), rays/s, the dominant kernel's
HBM roofline fraction and the CPU baseline (the oracle timed on this host's
cores, at 16 threads and at every core the process may use, on a bounded
sample of the same workload).  For cloud512 the line also carries the
equal-time relMSE of guided against plain rendering on the full frame,
against an FP64 converged render that CPU probe pixels accept.  Inside the timed steps every
kernel launch is bracketed by HIP events on the library's stream, which gives
each stage's average launch duration; the stage's algorithmic bytes come from
a separate counting run of the same frame (same launches, same paths).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"

# Algorithmic bytes per counted event: a trilinear density lookup reads 8 FP32
# corners, a majorant cell 4 B, a BVH node record 64 B (four quantized child
# boxes, child references and leaf counts), a primitive record 48 B
# (triangle: 3 x float4).  Ray-state traffic of the
# wavefront stages per ray (structure-of-arrays records):
#   trace  reads origin + direction (24 B), writes hit t / ref / kind (12 B)
#   medium reads origin, direction, hit t, pixel, sample, RNG count (40 B),
#          writes hit t, kind, RNG count (12 B)
#   shadow reads the 64 B queue record, read-modify-writes the 12 B radiance
#   shade  reads the 84 B path record, writes 64 B per continuation and per
#          shadow ray
# The film costs 24 B read + 24 B write per pixel (FP64 RGB), once per frame.
EV = {"density_lookups": 32, "majorant_cells": 4, "bvh_nodes": 64, "prim_tests": 48}
RAY_BYTES = {"trace": 36, "medium": 52, "shadow": 88}
SHADE_IN, SHADE_OUT = 84, 64


# a new path's record: origin, direction, throughput, radiance (48 B), pixel,
# sample, RNG count, depth (16 B), previous pdf (4 B)
REGEN_BYTES = 68


def stage_bytes(stage: str, c: dict, film_bytes: int, regen_in_medium: bool = False,
                shadow_in_shade: bool = False) -> int:
    """Algorithmic bytes a stage moves over one frame, from the event counters.
    regen_in_medium: volume-only scenes make their new paths inside the medium
    stage (no separate regen launch), whose bytes then count there.
    shadow_in_shade: volume-only plain renders track the shadow rays in the
    shading stage, which then moves their lookups and cells and no queue."""
    ext = {k: c[k] - c["shadow_" + k] for k in EV}
    sh = {k: c["shadow_" + k] for k in EV}
    if stage == "megakernel":
        return sum(c[k] * b for k, b in EV.items()) + film_bytes
    if stage == "trace":
        return ext["bvh_nodes"] * EV["bvh_nodes"] + ext["prim_tests"] * 48 + c["ext_rays"] * RAY_BYTES["trace"]
    if stage == "medium":
        return (ext["density_lookups"] * 32 + ext["majorant_cells"] * 4 + c["ext_rays"] * RAY_BYTES["medium"] +
                (c["paths"] * REGEN_BYTES if regen_in_medium else 0))
    if stage == "shadow":
        return sum(sh[k] * b for k, b in EV.items()) + c["shadow_rays"] * RAY_BYTES["shadow"]
    if stage == "shade":
        cont = max(c["ext_rays"] - c["paths"], 0)
        if shadow_in_shade:
            return c["ext_rays"] * SHADE_IN + cont * SHADE_OUT + sum(sh[k] * b for k, b in EV.items())
        return c["ext_rays"] * SHADE_IN + (cont + c["shadow_rays"]) * SHADE_OUT
    return 0


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    f = PROFILES / f"traffic_{kernel}_{workload}.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("TRAFFIC_BYTES_PER_LAUNCH="):
            return int(float(line.split("=", 1)[1]))
    return None


def batch_config(cfg) -> dict:
    """A batch configuration (tile_w, tile_h, samples_per_item, executor) for the JSON line."""
    if len(cfg) > 3 and cfg[3] == 1:
        return {"executor": "wavefront", "lanes": 4}
    return {"executor": "persistent", "item_shape": list(cfg[:3])}


def measured_valu(kernel: str, workload: str):
    """VALU issue fraction and active-lane fraction of `kernel` on `workload`
    from the committed rocprofv3 PMC summary, or None.  The hot stages are
    bound by vector-ALU issue, which is why their HBM fractions stay low."""
    f = PROFILES / "valu_utilization.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 3 and parts[0] == workload and parts[1].startswith(kernel + "<"):
            issue, lanes = parts[-2], parts[-1]
            return {"resource": "valu", "issue_frac": round(float(issue.rstrip("%")) / 100, 2),
                    "lanes_active_frac": round(float(lanes.rstrip("%")) / 100, 2),
                    "source": "profiles/valu_utilization.txt"}
    return None


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def host_cpu() -> dict:
    """The host's CPU model (/proc/cpuinfo, as lscpu reports it), the cores this
    process may run on (nproc) and the machine's logical CPU count."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"model": model, "nproc": nproc, "logical_cpus": os.cpu_count()}


def cpu_baseline(wl, budget_s: float):
    """Oracle (CPU restatement, FP32, same counter RNG and estimator) over the
    whole frame at the first k samples per pixel, k sized for about budget_s
    seconds: the same distribution of paths as the GPU frame (every pixel), so
    samples/s compare like for like.  A 1-sample pilot frame sets k.  Two legs:
    16 threads (OfflineEngine's pool, src/core/OfflineEngine.cpp:18) and every
    core the process may use (nproc); `value` is the nproc leg."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    host = host_cpu()
    t0 = time.perf_counter()
    w = O.World(wl.scene)
    build_s = time.perf_counter() - t0

    def run(k, offset, threads):
        t = time.perf_counter()
        w.render(wl.camera, wl.width, wl.height, k, wl.max_bounces, seed=1, sample_offset=offset, threads=threads)
        return time.perf_counter() - t, wl.width * wl.height * k

    legs = []
    for threads in sorted({16, host["nproc"]}):
        dt, n = run(1, 0, threads)
        k = int(max(1, min(wl.spp - 1, budget_s * (n / max(dt, 1e-9)) / (wl.width * wl.height))))
        dt, n = run(k, 1, threads)
        legs.append({"threads": threads, "value": n / dt, "paths": n, "seconds": round(dt, 2), "k": k})
        log(f"CPU baseline, {threads} threads: {n / dt / 1e6:.2f} M samples/s ({n} paths in {dt:.1f} s)")
    top = max(legs, key=lambda leg: leg["threads"])
    return {"value": top["value"], "unit": "samples/s", "cores": top["threads"], "kind": "port",
            "sample": f"whole {wl.width}x{wl.height} frame at samples 1..{top['k']} of {wl.spp} per pixel, "
                      f"{wl.max_bounces} bounces ({top['paths']} paths, {top['seconds']:.1f} s; 1-sample pilot frame "
                      f"and oracle scene build {build_s:.1f} s not timed)",
            "cpu_model": host["model"], "nproc": host["nproc"], "logical_cpus": host["logical_cpus"],
            "legs": [{"threads": leg["threads"], "value": round(leg["value"], 1), "paths": leg["paths"],
                      "seconds": leg["seconds"]} for leg in legs]}


def copy_calibration(nbytes: int, reps: int = 5):
    """Plain device-to-device copies of a known size, for scaling the rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (each copy reads and writes nbytes)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def bench_batch(args):
    """48 uneven volume jobs in one ne_batch_run.  The executor (persistent
    kernel with an item shape, or wavefront lanes) is autotuned first (untimed,
    on a 1/16-spp calibration run); a step is one run that renders every job
    at full spp.  The executor not chosen is timed beside it."""
    from narvalengine_amd import Batch, Context, configs
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL
    t0 = time.perf_counter()
    wls = configs.batch48()
    log(f"batch48: {len(wls)} jobs built in {time.perf_counter() - t0:.1f} s")
    ctx = Context(0)
    jobs = [dict(scene=ctx.upload(w.scene), camera=w.camera, width=w.width, height=w.height, spp=w.spp,
                 max_bounces=w.max_bounces, seed=1, job=k) for k, w in enumerate(wls)]
    b = Batch(ctx, jobs)
    cfg, tune_ms = b.autotune(spp_divisor=16)
    log(f"autotuned batch configuration {batch_config(cfg)} ({tune_ms:.1f} ms calibration run)")
    for i in range(args.warmup):
        b.run()
    paths = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    ms, job_ms = [], None
    ctx.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ms.append(b.run())
        job_ms = b.job_times()
        log(f"step {k + 1}/{args.steps}: launch {ms[-1]:.1f} ms")
    wall = time.perf_counter() - t_start
    value = args.steps * paths / wall
    launch_s = sum(ms) / len(ms) / 1e3
    persistent = batch_config(cfg)["executor"] == "persistent"
    # the other executor, timed once after a warm-up run
    executors = {batch_config(cfg)["executor"]: {"ms": round(launch_s * 1e3, 2), "batch_config": batch_config(cfg)}}
    others = [(8, 8, 4, 1)] if persistent else [(8, 8, 1, 0), (8, 8, 4, 0)]
    for oc in others:
        b.run(oc)
        t_o = b.run(oc)
        name = batch_config(oc)["executor"]
        if name not in executors or t_o < executors[name]["ms"]:
            executors[name] = {"ms": round(t_o, 2), "batch_config": batch_config(oc)}
    log(f"executors: {executors}")
    roofline, counters = None, None
    if not args.no_count:
        tot = {}
        for j in jobs:
            f = ctx.film(j["width"], j["height"])
            ctx.render(j["scene"], j["camera"], f, spp=j["spp"], max_bounces=j["max_bounces"], seed=1, job=j["job"],
                       mode=NE_MODE_MEGAKERNEL, count=True)
            for key, v in ctx.counters().items():
                tot[key] = tot.get(key, 0) + v
            f.close()
        counters = tot
        algo = sum(tot[k] * b_ for k, b_ in EV.items()) + 2 * 8 * 3 * sum(j["width"] * j["height"] for j in jobs)
        achieved = algo / launch_s / 1e9
        # persistent: the one k_batch launch; wavefront lanes: every stage
        # kernel of the run on all lanes together (the run's wall time)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": measured_traffic("k_batch", "batch48") if persistent else None,
                    "kernel": "k_batch" if persistent else "wavefront stages on 4 streams (whole run)",
                    "algorithmic_bytes_per_launch": int(algo), "avg_launch_ms": round(launch_s * 1e3, 3)}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle as O
        host = host_cpu()
        legs = []
        for threads in sorted({16, host["nproc"]}):
            n, dt = 0, 0.0
            for w in wls:  # every job, whole frame, first sample of each pixel
                world = O.World(w.scene)
                t = time.perf_counter()
                world.render(w.camera, w.width, w.height, 1, w.max_bounces, seed=1, threads=threads)
                dt += time.perf_counter() - t
                n += w.width * w.height
            legs.append({"threads": threads, "value": round(n / dt, 1), "paths": n, "seconds": round(dt, 2)})
        top = legs[-1]
        cpu = {"value": top["value"], "unit": "samples/s", "cores": top["threads"], "kind": "port",
               "sample": f"every job, whole frame at 1 sample per pixel ({top['paths']} paths, {top['seconds']:.1f} s)",
               "cpu_model": host["model"], "nproc": host["nproc"], "legs": legs}
    # the plain method beside it: the same jobs one after another through
    # ne_render (the default FP32 pipeline), each into its own film
    seq = None
    if not args.no_sequential:
        films = [ctx.film(j["width"], j["height"]) for j in jobs]
        for j, f in zip(jobs, films):  # untimed pass: pools and films allocated
            ctx.render(j["scene"], j["camera"], f, spp=j["spp"], max_bounces=j["max_bounces"], seed=1, job=j["job"],
                       sync=False)
        ctx.synchronize()
        for f in films:
            f.clear()
        done_ms = []
        t0 = time.perf_counter()
        for j, f in zip(jobs, films):
            ctx.render(j["scene"], j["camera"], f, spp=j["spp"], max_bounces=j["max_bounces"], seed=1, job=j["job"])
            done_ms.append((time.perf_counter() - t0) * 1e3)
        t_seq = time.perf_counter() - t0
        seq = {"method": "ne_render per job, one after another (FP32 wavefront)", "makespan_ms": round(t_seq * 1e3, 1),
               "value": round(paths / t_seq, 1), "unit": "samples/s",
               "batched_speedup": round(t_seq / (wall / args.steps), 3),
               "job_completion_ms": {"min": round(min(done_ms), 2), "median": round(float(np.median(done_ms)), 2),
                                     "max": round(max(done_ms), 2)}}
        log(f"sequential ne_render: {t_seq * 1e3:.1f} ms for all jobs ({paths / t_seq / 1e9:.3f} G samples/s)")
        for f in films:
            f.close()
    jm = np.asarray(job_ms, np.float64)
    out = {
        "metric": "samples/sec", "value": round(value, 1), "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural grids)",
        "config": {"workload": "batch48", "jobs": len(jobs), "paths_per_step": paths, "batch_config": batch_config(cfg),
                   "accumulation": "fixed-point i64"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "job_completion_ms": {"min": round(float(jm.min()), 2), "median": round(float(np.median(jm)), 2),
                              "max": round(float(jm.max()), 2), "all": [round(float(x), 2) for x in jm]},
        "executors": executors,
        "sequential": seq,
        "counters_per_step": counters,
    }
    print(json.dumps(out))


MFMA_BF16_PEAK_TFLOPS = 2500.0
GUIDE_FLOP_PER_QUERY = 2 * (35 * 64 + 3 * 64 * 64 + 64 * 40)  # 34,176


def rel_mse(x, ref, eps=1e-2):
    """mean over pixels and channels of (x - y)^2 / (y^2 + eps), linear radiance"""
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))


def probe_pixels(W: int, H: int, side: int = 16) -> np.ndarray:
    """The fixed probe set of the converged-reference check: a side x side grid
    over the cloud's part of the frame (x in [0.32, 0.68] W, y in [0.36, 0.66] H)."""
    xs = np.round(np.linspace(0.32 * W, 0.68 * W, side)).astype(np.int32)
    ys = np.round(np.linspace(0.36 * H, 0.66 * H, side)).astype(np.int32)
    return np.array([(x, y) for y in ys for x in xs], np.int32)


def converged_reference(ctx, g, wl, spp: int, passes: int, seed: int):
    """FP64 megakernel render at spp samples per pixel, in `passes` launches
    of spp / passes.  Returns the mean image, the variance of that mean per
    pixel and channel (from the spread of the passes, passes - 1 degrees of
    freedom) and the seconds taken."""
    from narvalengine_amd.abi import NE_FP64, NE_MODE_MEGAKERNEL
    W, H = wl.width, wl.height
    per = spp // passes
    f = ctx.film(W, H)
    s1 = np.zeros((H, W, 3))
    s2 = np.zeros((H, W, 3))
    prev = np.zeros((H, W, 3))
    t = time.perf_counter()
    for k in range(passes):
        ctx.render(g, wl.camera, f, spp=per, sample_offset=k * per, max_bounces=wl.max_bounces, seed=seed,
                   precision=NE_FP64, mode=NE_MODE_MEGAKERNEL)
        cur = f.read()
        img = (cur - prev) / per
        prev = cur
        s1 += img
        s2 += img * img
        if (k + 1) % 8 == 0 or k + 1 == passes:
            log(f"FP64 reference pass {k + 1}/{passes}")
    secs = time.perf_counter() - t
    f.close()
    mean = s1 / passes
    var_mean = np.maximum(s2 / passes - mean * mean, 0.0) / max(passes - 1, 1)
    return mean, var_mean, secs


def probe_check(wl, ref, ref_var, ref_dof: int, spp: int, threads: int, seed: int) -> dict:
    """Accepts the GPU FP64 reference only where the CPU oracle agrees: the
    oracle renders the probe pixels in FP64 at spp samples each (independent
    streams), and every probe and channel must agree with the reference within
    the Bonferroni bound of a 1% family-wise test: |z| below the Student-t
    quantile 1 - 0.01 / (2 n) with the reference's degrees of freedom (both
    estimates' variances from their own samples)."""
    from scipy.stats import t as student_t
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    W, H = wl.width, wl.height
    xy = probe_pixels(W, H)
    w = O.World(wl.scene, fp64=True)
    t = time.perf_counter()
    s1, s2 = w.render_pixels(wl.camera, W, H, xy, spp, wl.max_bounces, seed=seed, threads=threads)
    secs = time.perf_counter() - t
    m = s1 / spp
    var_m = np.maximum(s2 / spp - m * m, 0.0) / (spp - 1)
    r = ref[xy[:, 1], xy[:, 0]]
    rv = ref_var[xy[:, 1], xy[:, 0]]
    z = (r - m) / np.sqrt(np.maximum(var_m + rv, 1e-30))
    bound = float(student_t.ppf(1.0 - 0.01 / (2 * z.size), ref_dof))
    worst = np.argsort(-np.abs(z).reshape(-1))[:4]
    return {"probes": int(len(xy)), "spp": spp, "threads": threads, "seconds": round(secs, 2),
            "max_abs_z": round(float(np.abs(z).max()), 3), "z_bound": round(bound, 3),
            "mean_z2": round(float(np.mean(z * z)), 3), "count_abs_z_over_3": int((np.abs(z) > 3).sum()),
            "worst": [{"pixel": [int(xy[k // 3, 0]), int(xy[k // 3, 1])], "channel": int(k % 3),
                       "z": round(float(z.reshape(-1)[k]), 3), "reference": float(r.reshape(-1)[k]),
                       "cpu": float(m.reshape(-1)[k]), "reference_stderr": float(np.sqrt(rv.reshape(-1)[k])),
                       "cpu_stderr": float(np.sqrt(var_m.reshape(-1)[k]))} for k in worst],
            "mean_ratio_reference_to_cpu": round(float(r.sum() / max(m.sum(), 1e-30)), 5),
            "rel_stderr_cpu": round(float(np.sqrt(var_m.mean()) / max(m.mean(), 1e-30)), 6),
            "accepted": bool(np.abs(z).max() < bound)}


def relmse_parts(img, ref, ref_var, eps=1e-2):
    """relMSE of img against the reference, and the same with the reference's
    own noise (its per-pixel variance) taken out."""
    raw = rel_mse(img, ref, eps)
    noise = float(np.mean(ref_var / (ref ** 2 + eps)))
    return raw, raw - noise


def equal_time_relmse(ctx, g, wl, args, threads: int) -> dict:
    """Equal-time relMSE of guided against plain rendering on the full frame
    of the workload, against an FP64 converged render that the CPU probes
    accept.  For each budget of K frames: plain is the FP32 wavefront at K x
    the workload's spp, and sets the wall time.  Guided: a fresh guide trains
    online while it renders args.train_spp, then the frozen, tabulated guide
    renders the rest of the same wall time (training included; both phases'
    samples count), through the wavefront (guided) and through the persistent
    batch kernel (guided+batched).  The batch kernel's plain frame is timed
    beside them."""
    W, H, spp = wl.width, wl.height, wl.spp
    ref_spp = args.ref_mult * spp
    ref, ref_var, ref_s = converged_reference(ctx, g, wl, ref_spp, args.ref_passes, seed=1000)
    log(f"FP64 reference: {ref_spp} spp in {ref_s:.1f} s")
    probes = probe_check(wl, ref, ref_var, args.ref_passes - 1, args.probe_spp, threads, seed=2000)
    log(f"CPU probes: {probes}")
    f = ctx.film(W, H)
    ctx.render(g, wl.camera, f, spp=spp, seed=7, max_bounces=wl.max_bounces)  # warm
    out = {"frame": [W, H], "target": "guided+batched relMSE <= 0.5 x plain at equal time",
           "reference": {"kind": "FP64 megakernel", "spp": ref_spp, "passes": args.ref_passes,
                         "seconds": round(ref_s, 1), "relmse_noise": float(np.mean(ref_var / (ref ** 2 + 1e-2))),
                         "cpu_probes": probes},
           "budgets": []}
    for K in args.budgets:
        out["budgets"].append(equal_time_budget(ctx, g, wl, args, ref, ref_var, f, K))
    f.close()
    gb = [r["ratio_to_plain_time_normalised"] for b in out["budgets"] for r in b.get("guided_batched", [])]
    out["best_guided_batched_ratio"] = min(gb) if gb else None
    return out


def equal_time_budget(ctx, g, wl, args, ref, ref_var, f, K: int) -> dict:
    from narvalengine_amd import Batch, Guide
    W, H, spp = wl.width, wl.height, wl.spp
    # every wavefront render of the comparison on the pool the plain frame
    # gets by default (2^23 slots at this size), so that the pilot's time per
    # sample holds for the longer renders and no form runs on a smaller pool
    kw = dict(max_bounces=wl.max_bounces, pool_size=1 << 23)
    n_plain = K * spp
    f.clear()
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.render(g, wl.camera, f, spp=n_plain, seed=11, **kw)
    ctx.synchronize()
    t_plain = time.perf_counter() - t0
    e_plain, e_plain_c = relmse_parts(f.read() / n_plain, ref, ref_var)
    log(f"budget {K} frames: plain {n_plain} spp in {t_plain * 1e3:.1f} ms, relMSE {e_plain:.4e}")
    res = {"frames": K, "budget_ms": round(t_plain * 1e3, 2),
           "plain": {"spp": n_plain, "ms": round(t_plain * 1e3, 2), "relmse": e_plain,
                     "relmse_minus_ref_noise": e_plain_c},
           "guided": [], "guided_batched": []}

    def score(name, prob, n, t_all, t_train, steps, img, extra):
        e, e_c = relmse_parts(img, ref, ref_var)
        rec = {"guide_prob": prob, "guided_vertices": args.guided_vertices, "cache_res": args.guide_cache,
               "cache_dirs": args.guide_cache_dirs, "train_spp": args.train_spp, "train_pool": args.train_pool, "spp": n,
               "ms": round(t_all * 1e3, 2), "train_ms": round(t_train * 1e3, 2), "train_steps": steps,
               "relmse": e, "relmse_minus_ref_noise": e_c, "ratio_to_plain": round(e_c / e_plain_c, 4),
               # relMSE x time: the ratio at exactly equal time for an unbiased estimator
               "ratio_to_plain_time_normalised": round(e_c * t_all / (e_plain_c * t_plain), 4),
               "ratio_at_equal_spp": round(e_c * n / (e_plain_c * n_plain), 4), **extra}
        log(f"budget {K}, {name} p={prob}: {rec}")
        return rec

    # the batch kernel's plain frame
    job = dict(scene=g, camera=wl.camera, width=W, height=H, spp=n_plain, max_bounces=wl.max_bounces, seed=17, job=0)
    if args.batched:
        b = Batch(ctx, [job])
        cfg, _ = b.autotune(spp_divisor=16)
        b.run(cfg)
        ms = b.run(cfg)
        e, e_c = relmse_parts(b.read(0) / n_plain, ref, ref_var)
        b.close()
        res["plain_batched"] = {"spp": n_plain, "ms": round(ms, 2), "batch_config": batch_config(cfg), "relmse": e,
                                "relmse_minus_ref_noise": e_c,
                                "samples_per_s": round(W * H * n_plain / (ms * 1e-3), 1)}
        log(f"budget {K}: plain batched {res['plain_batched']}")
    for prob in args.guide_probs:
        gkw = dict(guided_vertices=args.guided_vertices, guide_prob=prob, cache_res=args.guide_cache,
                   cache_dirs=args.guide_cache_dirs)
        for form in ("guided", "guided_batched") if args.batched else ("guided",):
            guide = Guide(ctx, seed=22, **gkw)
            f.clear()
            ctx.synchronize()
            t0 = time.perf_counter()
            # a small path pool while training: each wavefront iteration ends
            # enough paths for one training step, and iterations are short
            ctx.render(g, wl.camera, f, spp=args.train_spp, seed=13, guide=guide, train=True,
                       **dict(kw, pool_size=args.train_pool))
            ctx.synchronize()
            t_train = time.perf_counter() - t0
            steps = guide.stats()["steps"]
            # frozen-guide samples (kept in the image) fill the rest of the
            # budget at the trained guide's own rate: a quarter-frame pilot
            # sets half of what is left, that render's rate sets the rest
            # (each render call has a fixed tail, which a short pilot
            # overstates per sample)
            done, acc = [0], [f.read()]

            def frozen(n, k):
                """n more frozen-guide samples (the k-th render); the timed seconds."""
                if form == "guided":
                    t1 = time.perf_counter()
                    ctx.render(g, wl.camera, f, spp=n, seed=13, sample_offset=args.train_spp + done[0], guide=guide,
                               **kw)
                    ctx.synchronize()
                    dt = time.perf_counter() - t1
                else:
                    # a first run untimed (graph capture, pools, tabulation), the second timed
                    bt = Batch(ctx, [dict(job, spp=n, seed=14 + k, guide=guide)])
                    bt.run(cfg)
                    ctx.synchronize()
                    t1 = time.perf_counter()
                    bt.run(cfg)
                    ctx.synchronize()
                    dt = time.perf_counter() - t1
                    acc[0] = acc[0] + bt.read(0)
                    bt.close()
                done[0] += n
                return dt

            n_pilot = max(1, spp // 4)
            dt = frozen(n_pilot, 0)
            spent, per = t_train + dt, dt / n_pilot
            n1 = int(0.5 * (t_plain - spent) / per)
            if n1 > 0:
                dt = frozen(n1, 1)
                spent, per = spent + dt, dt / n1
            n2 = int((t_plain - spent) / per)
            if n2 > 0:
                spent += frozen(n2, 2)
            img = (f.read() if form == "guided" else acc[0]) / (args.train_spp + done[0])
            extra = {("frozen_spp" if form == "guided" else "batched_spp"): done[0],
                     "renders": [n_pilot, max(n1, 0), max(n2, 0)]}
            if form != "guided":
                extra["batch_config"] = batch_config(cfg)
            res[form].append(score(form, prob, args.train_spp + done[0], spent, t_train, steps, img, extra))
            guide.close()
    return res


def pinhole_rays(cam, W, H, xy, n, rng):
    """n jittered camera rays per pixel (x, y) of a pinhole camera looking
    from cam.position at cam.look_at with vertical field of view cam.vfov."""
    pos = np.asarray(cam.position, np.float64)
    fwd = np.asarray(cam.look_at, np.float64) - pos
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(cam.up, np.float64))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = np.tan(np.radians(cam.vfov) / 2)
    px = np.repeat(xy[:, 0], n) + rng.random(len(xy) * n)
    py = np.repeat(xy[:, 1], n) + rng.random(len(xy) * n)
    u = (2 * px / W - 1) * th * (W / H)
    v = (1 - 2 * py / H) * th
    d = fwd[None] + u[:, None] * right[None] + v[:, None] * up[None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([np.tile(pos, (len(d), 1)), d], axis=1).astype(np.float32)


def bench_decompose(args):
    """Where the plain estimator's variance comes from on the cloud.  Camera
    rays through the probe pixels (cloud512's 16 x 16 grid over the cloud),
    args.decompose_spp jittered rays each, traced by Integrator::Li (ne_li)
    with identical streams in five scene variants, split every sample into
        escape  the camera ray leaves without a collision (sky seen directly)
        v1      the first scattering vertex: its light samples and the MIS
                environment hit of its continuation
        rest    everything after the phase continuation of the first vertex
    and v1 / rest further into the sun's and the sky's share.  Each share's
    per-sample variance is normalised as relMSE is, Var / (mean^2 + 0.01), and
    averaged over the pixels; a directional proposal at the scattering
    vertices only changes the `rest` share."""
    from narvalengine_amd import Context, configs
    wl = configs.cloud512()
    W, H = wl.width, wl.height
    xy = probe_pixels(W, H)
    n = args.decompose_spp
    rays = pinhole_rays(wl.camera, W, H, xy, n, np.random.default_rng(5))
    ctx = Context(0)

    def li(scene, bounces):
        g = ctx.upload(scene)
        out = ctx.li(g, rays, bounces, seed=77).astype(np.float64).reshape(len(xy), n, 3)
        g.close()
        return out

    full = li(wl.scene, wl.max_bounces)
    b0 = li(wl.scene, 0)
    b1 = li(wl.scene, 1)
    sun_only = configs.cloud512().scene
    sun_only.set_env(sun_only.env["kind"], c0=(0, 0, 0), c1=(0, 0, 0), c2=(0, 0, 0))
    sky_only = configs.cloud512().scene
    sky_only.set_sun(direction=sky_only.sun["dir"], radiance=(0.0, 0.0, 0.0))
    sun_full, sun_b1 = li(sun_only, wl.max_bounces), li(sun_only, 1)
    sky_full, sky_b1 = li(sky_only, wl.max_bounces), li(sky_only, 1)
    parts = {"escape": b0, "v1": b1 - b0, "rest": full - b1,
             "v1_sun": sun_b1, "v1_sky": sky_b1 - b0, "rest_sun": sun_full - sun_b1, "rest_sky": sky_full - sky_b1}
    mean = full.mean(axis=1)
    norm = mean ** 2 + 1e-2

    def relvar(x):
        return float(np.mean(x.var(axis=1) / norm))

    def relcov(x, y):
        return float(np.mean(((x - x.mean(axis=1, keepdims=True)) * (y - y.mean(axis=1, keepdims=True))).mean(axis=1)
                             / norm))

    total = relvar(full)
    out = {"metric": "per-sample variance shares of the plain estimator (cloud512 probe pixels)",
           "pixels": int(len(xy)), "samples_per_pixel": n, "relvar_total": total,
           "relvar": {k: relvar(v) for k, v in parts.items()},
           "share": {k: round(relvar(v) / total, 4) for k, v in parts.items()},
           "relcov": {"escape,v1": relcov(parts["escape"], parts["v1"]),
                      "escape,rest": relcov(parts["escape"], parts["rest"]),
                      "v1,rest": relcov(parts["v1"], parts["rest"])},
           "check_sum_of_parts": round(float(np.abs(b0 + parts["v1"] + parts["rest"] - full).max()), 9),
           "check_sun_plus_sky": round(float(np.abs(sun_full + sky_full - full).max()), 5),
           "mean_radiance": [round(float(x), 5) for x in mean.mean(axis=0)],
           "fraction_escaping": round(float((b0.sum(axis=2) > 0).mean()), 4)}
    print(json.dumps(out))


def bench_guide_sweep(args):
    """Per-sample value of the learned proposal on the cloud at a reduced frame
    (the cloud512 scene, camera and 128 bounces at args.sweep_width x
    args.sweep_height).  For each (guided vertices, proposal probability): a
    fresh guide trains online over args.sweep_train_spp, then frozen renders
    of 256 spp (4 seeds) are compared with plain renders of 256 spp against an
    FP64 render at args.sweep_ref_spp.  Reports relMSE at equal spp, the time
    per frame, and their product (inverse efficiency) relative to plain."""
    from narvalengine_amd import Context, Guide, configs
    from narvalengine_amd.abi import NE_FP64, NE_MODE_MEGAKERNEL
    W, H = args.sweep_width, args.sweep_height
    wl = configs.cloud512(width=W, height=H)
    ctx = Context(0)
    g = ctx.upload(wl.scene)
    kw = dict(max_bounces=wl.max_bounces)
    f = ctx.film(W, H)
    t = time.perf_counter()
    ctx.render(g, wl.camera, f, spp=args.sweep_ref_spp, seed=1000, precision=NE_FP64, mode=NE_MODE_MEGAKERNEL, **kw)
    ref = f.read() / args.sweep_ref_spp
    log(f"reference {args.sweep_ref_spp} spp FP64 in {time.perf_counter() - t:.1f} s")

    def frames(n_seeds, **extra):
        errs, ms = [], []
        for k in range(n_seeds):
            f.clear()
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.render(g, wl.camera, f, spp=256, seed=50 + k, **kw, **extra)
            ms.append((time.perf_counter() - t0) * 1e3)
            errs.append(rel_mse(f.read() / 256, ref))
        return float(np.mean(errs)), float(np.median(ms))

    frames(1)
    e_plain, ms_plain = frames(4)
    log(f"plain: relMSE {e_plain:.4e}, {ms_plain:.2f} ms per frame")
    rows = []
    for spec in args.sweep_configs.split(","):
        parts = spec.split(":")
        v, p = parts[0], parts[1]
        nc = int(parts[2]) if len(parts) > 2 else 0
        guide = Guide(ctx, seed=31, guided_vertices=int(v), guide_prob=float(p), cache_res=nc,
                      cache_dirs=args.guide_cache_dirs)
        f.clear()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.render(g, wl.camera, f, spp=args.sweep_train_spp, seed=7, guide=guide, train=True, **kw)
        ctx.synchronize()
        t_train = time.perf_counter() - t0
        frames(1, guide=guide)
        e, ms = frames(4, guide=guide)
        st = guide.stats()
        guide.close()
        row = {"guided_vertices": int(v), "guide_prob": float(p), "cache_res": nc, "train_steps": st["steps"],
               "train_s": round(t_train, 2), "relmse": e, "ms_per_frame": round(ms, 3),
               "relmse_ratio_equal_spp": round(e / e_plain, 4), "time_ratio": round(ms / ms_plain, 4),
               "inverse_efficiency_ratio": round(e * ms / (e_plain * ms_plain), 4)}
        rows.append(row)
        log(f"guided {spec}: {row}")
    print(json.dumps({"metric": "learned proposal vs phase sampling (cloud, reduced frame)", "frame": [W, H],
                      "reference_spp": args.sweep_ref_spp, "train_spp": args.sweep_train_spp,
                      "plain": {"relmse": e_plain, "ms_per_frame": round(ms_plain, 3)}, "guided": rows}))


def bench_quality(args):
    """Image error of each image workload: relMSE of the plain GPU render
    (FP32 wavefront, the workload's spp) against an FP64 converged render of
    the same scene (FP64 megakernel at ref_mult x spp; the FP64 path is
    pinned to the oracle's FP64 integrator by the parity tests), and a CPU
    cross-check: the oracle's own FP64 render of a 16 x 16 pixel block
    against the same block of the converged render."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    from narvalengine_amd import Context, configs
    from narvalengine_amd.abi import NE_FP64, NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    ctx = Context(0)
    threads = max(1, min(16, os.cpu_count() or 1))
    results = {}
    for name in ("volume128", "surface1m", "cloud512"):
        wl = configs.CONFIGS[name]()
        g = ctx.upload(wl.scene)
        W, H = wl.width, wl.height
        f = ctx.film(W, H)
        ctx.render(g, wl.camera, f, spp=wl.spp, max_bounces=wl.max_bounces, seed=1, mode=NE_MODE_WAVEFRONT)
        img = f.read() / wl.spp
        f.close()
        ref_spp = wl.spp * args.ref_mult
        fr = ctx.film(W, H)
        t = time.perf_counter()
        for k in range(args.ref_mult):  # in launches of the workload's spp (progress lines in between)
            ctx.render(g, wl.camera, fr, spp=wl.spp, sample_offset=k * wl.spp, max_bounces=wl.max_bounces, seed=2,
                       precision=NE_FP64, mode=NE_MODE_MEGAKERNEL)
            log(f"{name}: FP64 reference {k + 1}/{args.ref_mult}")
        t_ref = time.perf_counter() - t
        ref = fr.read() / ref_spp
        fr.close()
        e = rel_mse(img, ref)
        # the reference's own noise adds spp / ref_spp of the image's error
        e_corr = e / (1.0 + wl.spp / ref_spp)
        x0, y0 = W // 2 - 8, H // 2 - 8
        rect = (x0, y0, x0 + 16, y0 + 16)
        t = time.perf_counter()
        cpu = O.World(wl.scene, fp64=True).render(wl.camera, W, H, args.cpu_block_spp, wl.max_bounces, rect=rect,
                                                   seed=3, threads=threads)
        t_cpu = time.perf_counter() - t
        blk = cpu[y0:y0 + 16, x0:x0 + 16] / args.cpu_block_spp
        rblk = ref[y0:y0 + 16, x0:x0 + 16]
        results[name] = {"spp": wl.spp, "reference_spp": ref_spp, "relmse": e, "relmse_minus_reference_noise": e_corr,
                         "reference_seconds": round(t_ref, 2),
                         "cpu_fp64_block": {"rect": list(rect), "spp": args.cpu_block_spp,
                                            "mean_ratio_to_reference": float(blk.sum() / max(rblk.sum(), 1e-30)),
                                            "relmse_to_reference": rel_mse(blk, rblk), "seconds": round(t_cpu, 2),
                                            "threads": threads}}
        log(f"{name}: relMSE {e:.3e} (reference {ref_spp} spp FP64, {t_ref:.1f} s), CPU block ratio "
            f"{results[name]['cpu_fp64_block']['mean_ratio_to_reference']:.4f}")
        g.close()
    print(json.dumps({"metric": "relMSE vs FP64 converged render", "unit": "relMSE (mean (x-y)^2/(y^2+0.01))",
                      "n_gpus": 1, "higher_is_better": False, "dtype": "f32 (reference f64)",
                      "data": "synthetic (seeded procedural scenes)", "ref_mult": args.ref_mult,
                      "workloads": results}))


def bench_guided(args):
    """Learned proposal vs plain HG sampling at equal wall time on the 512^3
    cloud (the cloud512 scene, camera and bounce limit at 480x270): relMSE of
    each against a 16384-spp reference.  The guided run trains online during the
    render; its spp are chosen from a pilot so that its wall time, training
    included, matches the plain run's."""
    from narvalengine_amd import Context, Guide, configs
    W, H = 480, 270
    wl = configs.cloud512(width=W, height=H)
    ctx = Context(0)
    g_s = ctx.upload(wl.scene)
    kw = dict(max_bounces=wl.max_bounces)
    ref_spp, plain_spp = 16384, args.plain_spp
    gkw = dict(guided_vertices=args.guided_vertices, guide_prob=args.guide_prob)

    def timed_render(spp, seed, **extra):
        f = ctx.film(W, H)
        ctx.synchronize()
        t = time.perf_counter()
        ctx.render(g_s, wl.camera, f, spp=spp, seed=seed, **kw, **extra)
        dt = time.perf_counter() - t
        img = f.read() / spp
        f.close()
        return img, dt

    timed_render(64, 7)  # warm-up
    ref, t_ref = timed_render(ref_spp, 1000)
    log(f"reference: {ref_spp} spp in {t_ref:.2f} s")
    plain, t_plain = timed_render(plain_spp, 1)
    log(f"plain: {plain_spp} spp in {t_plain * 1e3:.1f} ms")
    pilot = Guide(ctx, seed=11, **gkw)
    timed_render(256, 2, guide=pilot, train=True)  # first guided render allocates the record pool
    _, t_pilot = timed_render(256, 5, guide=pilot, train=True)
    pilot.close()
    guided_spp = max(16, int(256 * t_plain / t_pilot))  # same wall time as the plain run
    guide = Guide(ctx, seed=12, **gkw)
    guided, t_guided = timed_render(guided_spp, 3, guide=guide, train=True)
    st = guide.stats()
    log(f"guided: {guided_spp} spp in {t_guided * 1e3:.1f} ms, guide {st}")
    # stage times and event counts of a plain and a guided frame (trained
    # guide, no training) at 256 spp
    f = ctx.film(W, H)
    ctx.render(g_s, wl.camera, f, spp=256, seed=4, stage_timed=True, **kw)
    plain_stages = ctx.stage_times()
    ctx.render(g_s, wl.camera, f, spp=256, seed=4, count=True, **kw)
    plain_cnt = ctx.counters()
    ctx.render(g_s, wl.camera, f, spp=256, seed=4, guide=guide, stage_timed=True, **kw)
    stages = ctx.stage_times()
    ctx.render(g_s, wl.camera, f, spp=256, seed=4, guide=guide, count=True, **kw)
    cnt = ctx.counters()
    g_ms, g_n = stages["guide"]
    flops = cnt["guide_queries"] * GUIDE_FLOP_PER_QUERY
    achieved = flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 6), "traffic": None, "kernel": "k_wf_guide",
                "flop_per_launch": int(flops / max(g_n, 1)), "avg_launch_ms": round(g_ms / max(g_n, 1), 4)}
    e_plain, e_guided = rel_mse(plain, ref), rel_mse(guided, ref)
    # per-sample efficiency: relMSE scales as 1 / spp for an unbiased estimator
    per_sample = (e_guided * guided_spp) / (e_plain * plain_spp)
    out = {
        "metric": "relMSE at equal time (guided / plain)", "value": round(e_guided / e_plain, 4), "unit": "ratio",
        "n_gpus": 1, "higher_is_better": False, "dtype": "f32 (MLP bf16 / fp32 acc)",
        "data": "synthetic (seeded procedural grid)",
        "config": {"workload": "cloud512_guided", "width": W, "height": H, "max_bounces": wl.max_bounces,
                   "reference_spp": ref_spp, "plain_spp": plain_spp, "guided_spp": guided_spp,
                   "guided_vertices": args.guided_vertices, "guide_prob": args.guide_prob},
        "relmse_plain": e_plain, "relmse_guided": e_guided, "relmse_ratio_at_equal_spp": round(per_sample, 4),
        "ms_plain": round(t_plain * 1e3, 2), "ms_guided_incl_training": round(t_guided * 1e3, 2),
        "guide": st, "guided_stage_ms": {k: round(v[0], 2) for k, v in stages.items() if v[1]},
        "plain_stage_ms": {k: round(v[0], 2) for k, v in plain_stages.items() if v[1]},
        "ext_rays_per_frame": {"plain": plain_cnt["ext_rays"], "guided": cnt["ext_rays"]},
        "density_lookups_per_frame": {"plain": plain_cnt["density_lookups"], "guided": cnt["density_lookups"]},
        "guide_queries_per_frame": cnt["guide_queries"],
        "roofline": roofline,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cloud512",
                    choices=["surface1m", "volume128", "cloud512", "batch48", "guided", "quality", "decompose", "guide-sweep"])
    ap.add_argument("--mode", default="auto", choices=["auto", "megakernel", "wavefront"])
    ap.add_argument("--pool", type=int, default=0, help="wavefront path slots (0 = library default)")
    ap.add_argument("--sort", action="store_true", help="wavefront: shade chunks in (event kind, material) order")
    ap.add_argument("--spatial", action="store_true",
                    help="wavefront: tracking stages take each chunk's rays in Morton order of their origins")
    ap.add_argument("--maj-res", type=int, default=0, help="majorant cells per axis (0 = the workload's)")
    ap.add_argument("--no-empty-skip", action="store_true",
                    help="majorant DDA reads every cell (no skipping of empty runs; same image)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--no-sequential", action="store_true", help="batch48: skip the one-job-after-another leg")
    ap.add_argument("--samples-per-thread", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--plain-spp", type=int, default=1024, help="guided workload: plain run's spp (sets the time)")
    ap.add_argument("--ref-mult", type=int, default=64,
                    help="reference spp = ref_mult x the workload's spp (quality workload and equal-time block)")
    ap.add_argument("--cpu-block-spp", type=int, default=1024, help="quality workload: oracle FP64 block spp")
    ap.add_argument("--guided-vertices", type=int, default=1)
    ap.add_argument("--guide-prob", type=float, default=0.25)
    ap.add_argument("--decompose-spp", type=int, default=1024, help="decompose workload: rays per probe pixel")
    ap.add_argument("--sweep-width", type=int, default=240)
    ap.add_argument("--sweep-height", type=int, default=135)
    ap.add_argument("--sweep-ref-spp", type=int, default=16384)
    ap.add_argument("--sweep-train-spp", type=int, default=4096)
    ap.add_argument("--sweep-configs", default="1:0.25,1:0.5,4:0.25,4:0.5,128:0.1,128:0.25,128:0.5,"
                                                 "1:0.5:16,4:0.5:16,128:0.25:16,128:0.5:16",
                    help="guide-sweep workload: guided_vertices:guide_prob[:cache_res] triples "
                         "(cache_res 0 or absent: per-vertex network)")
    ap.add_argument("--guide-cache", type=int, default=16, help="equal-time block: tabulation cells per axis")
    ap.add_argument("--guide-cache-dirs", type=int, default=4, help="tabulation direction bins per axis")
    ap.add_argument("--no-quality", action="store_true",
                    help="cloud512: skip the equal-time relMSE block (FP64 reference, CPU probes, guided runs)")
    ap.add_argument("--ref-passes", type=int, default=64,
                    help="equal-time block: the FP64 reference in this many launches (its variance estimate)")
    ap.add_argument("--probe-spp", type=int, default=65536, help="equal-time block: CPU FP64 spp per probe pixel")
    ap.add_argument("--train-spp", type=int, default=8, help="equal-time block: spp rendered while the guide trains")
    ap.add_argument("--train-pool", type=int, default=1 << 17,
                    help="equal-time block: wavefront path slots of the training render")
    ap.add_argument("--budgets", default="1,16",
                    help="equal-time block: time budgets in frames of the workload's spp (plain wavefront time)")
    ap.add_argument("--no-batched", dest="batched", action="store_false",
                    help="equal-time block: skip the batch-kernel legs (plain batched, guided+batched)")
    ap.add_argument("--guide-probs", default="0.5,0.25",
                    help="equal-time block: one guided run per proposal probability")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run plain copies of BYTES (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    args.guide_probs = [float(x) for x in args.guide_probs.split(",") if x]
    args.budgets = [int(x) for x in args.budgets.split(",") if x]
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the path shards by "
              "pixel tiles / sample ranges with a final FP64 reduce, which is not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.workload == "batch48":
        bench_batch(args)
        return
    if args.workload == "guided":
        bench_guided(args)
        return
    if args.workload == "quality":
        bench_quality(args)
        return
    if args.workload == "decompose":
        bench_decompose(args)
        return
    if args.workload == "guide-sweep":
        bench_guide_sweep(args)
        return

    from narvalengine_amd import Context, configs
    from narvalengine_amd.abi import NE_MODE_AUTO, NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    mode = {"auto": NE_MODE_AUTO, "megakernel": NE_MODE_MEGAKERNEL, "wavefront": NE_MODE_WAVEFRONT}[args.mode]
    t0 = time.perf_counter()
    wl = configs.CONFIGS[args.workload]()
    if args.maj_res and wl.scene.volume is not None:
        wl.scene.volume["maj_res"] = args.maj_res
    if args.no_empty_skip and wl.scene.volume is not None:
        wl.scene.volume["empty_skip"] = False
    log(f"workload {wl.name}: {wl.width}x{wl.height} spp {wl.spp} bounces {wl.max_bounces} "
        f"(scene built in {time.perf_counter() - t0:.1f} s)")
    ctx = Context(0)
    g = ctx.upload(wl.scene)
    info = g.info()
    log(f"scene on device: {info}")
    film = ctx.film(wl.width, wl.height)
    kw = dict(spp=wl.spp, max_bounces=wl.max_bounces, seed=1, samples_per_thread=args.samples_per_thread,
              block_size=args.block, mode=mode, pool_size=args.pool, sort=args.sort, spatial=args.spatial)
    for i in range(args.warmup):
        ctx.render(g, wl.camera, film, sample_offset=(1000 + i) * wl.spp, **kw)
        log(f"warmup {i + 1}/{args.warmup} done")
    ctx.synchronize()
    # timed region: K full frames; every kernel launch is bracketed by HIP
    # events on the context stream (per-stage durations)
    stage_ms, stage_n = {}, {}
    t_start = time.perf_counter()
    for k in range(args.steps):
        ctx.render(g, wl.camera, film, sample_offset=k * wl.spp, sync=False, stage_timed=True, **kw)
        st = ctx.stage_times()
        for name, (ms, n) in st.items():
            stage_ms[name] = stage_ms.get(name, 0.0) + ms
            stage_n[name] = stage_n.get(name, 0) + n
        log(f"step {k + 1}/{args.steps}: " + ", ".join(f"{n} {ms:.1f} ms" for n, (ms, c) in st.items() if c))
    ctx.synchronize()
    wall = time.perf_counter() - t_start
    paths_per_step = wl.width * wl.height * wl.spp
    value = args.steps * paths_per_step / wall
    step_s = wall / args.steps

    roofline, rays, counters, stages = None, None, None, None
    dom = max(stage_ms, key=lambda n: stage_ms[n])
    if not args.no_count:
        cfilm = ctx.film(wl.width, wl.height)
        ctx.render(g, wl.camera, cfilm, count=True, **kw)
        counters = ctx.counters()
        log(f"counters: {counters}")
        rays = counters["ext_rays"] + counters["shadow_rays"]
        film_bytes = 48 * wl.width * wl.height
        stages = {}
        for name in stage_ms:
            if not stage_n[name]:
                continue
            per_launch_ms = stage_ms[name] / stage_n[name]
            launches_per_frame = stage_n[name] / args.steps
            b = stage_bytes(name, counters, film_bytes, regen_in_medium=not stage_n.get("regen"),
                            shadow_in_shade=not stage_n.get("shadow") and counters["shadow_rays"] > 0)
            b /= launches_per_frame
            stages[name] = {"ms_per_frame": round(stage_ms[name] / args.steps, 3),
                            "launches_per_frame": launches_per_frame,
                            "avg_launch_ms": round(per_launch_ms, 4),
                            "algorithmic_bytes_per_launch": int(b),
                            "achieved_GBs": round(b / (per_launch_ms * 1e-3) / 1e9, 2) if b else None}
        d = stages[dom]
        kernel = {"megakernel": "k_render", "trace": "k_wf_trace", "medium": "k_wf_medium", "shade": "k_wf_shade",
                  "shadow": "k_wf_shadow", "plan": "k_wf_plan", "regen": "k_wf_regen"}.get(dom, dom)
        achieved = d["achieved_GBs"] or 0.0
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": measured_traffic(kernel, wl.name),
                    "kernel": kernel, "algorithmic_bytes_per_launch": d["algorithmic_bytes_per_launch"],
                    "avg_launch_ms": d["avg_launch_ms"], "limiter": measured_valu(kernel, wl.name)}
    cpu = None if args.no_cpu else cpu_baseline(wl, args.cpu_seconds)
    quality = None
    if wl.name == "cloud512" and not args.no_quality:
        quality = equal_time_relmse(ctx, g, wl, args, threads=host_cpu()["nproc"])
    out = {
        "metric": "samples/sec", "value": round(value, 1), "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": wl.name, "width": wl.width, "height": wl.height, "spp": wl.spp,
                   "max_bounces": wl.max_bounces, "paths_per_step": paths_per_step, "accumulation": "f64",
                   "mode": args.mode, "pool": args.pool, "material_sort": args.sort, "spatial_order": args.spatial,
                   "maj_res": wl.scene.volume["maj_res"] if wl.scene.volume is not None else None,
                   "empty_skip": (wl.scene.volume.get("empty_skip", True) if wl.scene.volume is not None
                                  else None)},
        "rays_per_sec": round(rays / step_s, 1) if rays else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "gpu_vs_cpu_legs": {str(leg["threads"]): round(value / leg["value"], 1) for leg in cpu["legs"]} if cpu else None,
        "equal_time_relmse": quality,
        "stages": stages,
        "counters_per_step": counters,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
