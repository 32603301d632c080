"""Context / scene / film handles over the C ABI (include/ne_gpu.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import (BatchConfig, Counters, GuideConfig, GuideStats, Job, NarvalError, RenderParams, SceneInfo, check, lib, NE_FP32,
                  NE_FLAG_COUNT, NE_FLAG_TIME, NE_FLAG_STAGE_TIME, NE_FLAG_SORT, NE_FLAG_FIXED_FILM, NE_FLAG_SPATIAL,
                  NE_FLAG_WIDE_OFFSETS, NE_FLAG_TEST_PLAN_STALL, STAGES)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Context:
    """One HIP device + stream (ne_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        check(self._lib.ne_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self._lib.ne_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int):
        check(st, self.h)

    def synchronize(self):
        self.check(self._lib.ne_synchronize(self.h))

    def upload(self, scene) -> "GpuScene":
        return GpuScene(self, scene)

    def film(self, width: int, height: int) -> "Film":
        return Film(self, width, height)

    def render(self, gscene: "GpuScene", camera, film: "Film", *, spp: int, max_bounces: int, seed: int = 0,
               sample_offset: int = 0, job: int = 0, rect=None, precision: int = NE_FP32, count: bool = False,
               samples_per_thread: int = 0, block_size: int = 0, sync: bool = True, timed: bool = False,
               stage_timed: bool = False, guide=None, train: bool = False,
               occupancy: int = 0, mode: int = 0, pool_size: int = 0, sort: bool = False,
               fixed_film: bool = False, noalloc: bool = False, spatial: bool = False, wide: bool = False,
               plan_stall: bool = False):
        p = RenderParams()
        p.width, p.height = film.width, film.height
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, film.width, film.height)
        p.x0, p.y0, p.x1, p.y1 = x0, y0, x1, y1
        p.spp, p.max_bounces, p.seed = spp, max_bounces, seed
        p.sample_offset, p.job, p.precision = sample_offset, job, precision
        p.flags = (NE_FLAG_COUNT if count else 0) | (NE_FLAG_TIME if timed else 0) | \
            (NE_FLAG_STAGE_TIME if stage_timed else 0) | (NE_FLAG_SORT if sort else 0) | \
            (NE_FLAG_FIXED_FILM if fixed_film else 0) | (NE_FLAG_SPATIAL if spatial else 0) | \
            (NE_FLAG_WIDE_OFFSETS if wide else 0) | (NE_FLAG_TEST_PLAN_STALL if plan_stall else 0)
        p.samples_per_thread, p.block_size, p.occupancy = samples_per_thread, block_size, occupancy
        p.mode, p.pool_size = mode, pool_size
        cam = camera.desc()
        if noalloc:  # megakernel with its partial sums in the static device scratch
            self.check(self._lib.ne_render_noalloc(self.h, gscene.h, C.byref(cam), C.byref(p), film.h))
        elif guide is not None:
            self.check(self._lib.ne_render_guided(self.h, gscene.h, C.byref(cam), C.byref(p), film.h, guide.h,
                                                  1 if train else 0))
        else:
            self.check(self._lib.ne_render(self.h, gscene.h, C.byref(cam), C.byref(p), film.h))
        if sync:
            self.synchronize()

    def kernel_ms(self) -> float:
        ms = C.c_float()
        self.check(self._lib.ne_get_kernel_time(self.h, C.byref(ms)))
        return float(ms.value)

    def stage_times(self) -> dict:
        """{stage: (total ms, launches)} of the last render issued with stage_timed=True."""
        ms = (C.c_float * len(STAGES))()
        n = (C.c_uint32 * len(STAGES))()
        self.check(self._lib.ne_get_stage_times(self.h, ms, n))
        return {name: (float(ms[k]), int(n[k])) for k, name in enumerate(STAGES)}

    def counters(self) -> dict:
        c = Counters()
        self.check(self._lib.ne_get_counters(self.h, C.byref(c)))
        return c.as_dict()

    def li(self, gscene: "GpuScene", rays: np.ndarray, max_bounces: int, seed: int = 0, sample: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = np.zeros((len(rays), 3), np.float32)
        self.check(self._lib.ne_li(self.h, gscene.h, len(rays), _fp(rays), max_bounces, seed, sample, _fp(out)))
        return out

    # ---- parity entry points
    def test_primitive(self, kind: int, rays: np.ndarray, prims: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        prims = np.ascontiguousarray(prims, np.float32)
        n = len(rays)
        hit = np.zeros(n, np.int32)
        t0 = np.zeros(n, np.float32)
        t1 = np.zeros(n, np.float32)
        self.check(self._lib.ne_test_primitive(self.h, kind, n, _fp(rays), _fp(prims), _ip(hit), _fp(t0), _fp(t1)))
        return hit, t0, t1

    def test_closest(self, gscene, rays, tmin=1e-4, tmax=np.inf):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        t = np.zeros(n, np.float32)
        ty = np.zeros(n, np.int32)
        ix = np.zeros(n, np.int32)
        self.check(self._lib.ne_test_closest(self.h, gscene.h, n, _fp(rays), tmin, tmax, _fp(t), _ip(ty), _ip(ix)))
        return t, ty, ix

    def test_transmittance(self, gscene, rays, tmax, seed=0, precision=NE_FP32):
        rays = np.ascontiguousarray(rays, np.float32)
        tmax = np.ascontiguousarray(tmax, np.float32)
        out = np.zeros(len(rays), np.float64)
        self.check(self._lib.ne_test_transmittance(self.h, gscene.h, len(rays), _fp(rays), _fp(tmax), seed, precision,
                                                   _dp(out)))
        return out

    def test_delta_track(self, gscene, rays, seed=0, precision=NE_FP32):
        rays = np.ascontiguousarray(rays, np.float32)
        sc = np.zeros(len(rays), np.int32)
        t = np.zeros(len(rays), np.float64)
        self.check(self._lib.ne_test_delta_track(self.h, gscene.h, len(rays), _fp(rays), seed, precision, _ip(sc),
                                                 _dp(t)))
        return sc, t

    def test_density(self, gscene, gcs):
        gcs = np.ascontiguousarray(gcs, np.float32)
        out = np.zeros(len(gcs), np.float32)
        self.check(self._lib.ne_test_density(self.h, gscene.h, len(gcs), _fp(gcs), _fp(out)))
        return out

    def test_majorants(self, gscene, maj_res: int):
        out = np.zeros((maj_res, maj_res, maj_res), np.float32)
        self.check(self._lib.ne_test_majorants(self.h, gscene.h, _fp(out), out.size))
        return out

    def test_bsdf(self, kind: int, n: int, ins, out_width: int):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in ins] + [None] * (4 - len(ins))
        out = np.zeros((n, out_width), np.float32)
        ptrs = [C.cast(None, C.POINTER(C.c_float)) if a is None else _fp(a) for a in arrs]
        self.check(self._lib.ne_test_bsdf(self.h, kind, n, *ptrs, _fp(out)))
        return out

    def test_rng(self, seed, pixel, sample, job, n):
        out = np.zeros(n, np.float32)
        self.check(self._lib.ne_test_rng(self.h, seed, pixel, sample, job, n, _fp(out)))
        return out


class GpuScene:
    """A scene resident in HBM (ne_scene)."""

    def __init__(self, ctx: Context, scene):
        self.ctx = ctx
        desc, keep = scene.desc()
        h = C.c_void_p()
        ctx.check(ctx._lib.ne_scene_upload(ctx.h, C.byref(desc), C.byref(h)))
        self.h = h

    def info(self) -> dict:
        i = SceneInfo()
        self.ctx.check(self.ctx._lib.ne_scene_get_info(self.h, C.byref(i)))
        return {n: getattr(i, n) for n, _ in i._fields_ if n != "pad"}

    def close(self):
        if self.h:
            self.ctx._lib.ne_scene_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Film:
    """FP64 per-pixel radiance sums in HBM (ne_film)."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx, self.width, self.height = ctx, width, height
        h = C.c_void_p()
        ctx.check(ctx._lib.ne_film_create(ctx.h, width, height, C.byref(h)))
        self.h = h

    def clear(self):
        self.ctx.check(self.ctx._lib.ne_film_clear(self.h))

    def read(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), np.float64)
        self.ctx.check(self.ctx._lib.ne_film_read(self.h, _dp(out)))
        return out

    def tonemap(self, spp: int) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), np.float32)
        self.ctx.check(self.ctx._lib.ne_film_tonemap(self.h, spp, _fp(out)))
        return out

    def close(self):
        if self.h:
            self.ctx._lib.ne_film_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Context", "GpuScene", "Film", "NarvalError"]


class Batch:
    """Many render jobs in one call (ne_batch).  jobs: list of dicts with
    scene (GpuScene), camera, width, height, spp, max_bounces and optional
    seed / job / guide.  A configuration is (tile_w, tile_h, samples_per_item)
    for the persistent kernel, or (tile_w, tile_h, samples_per_item, executor)
    with executor NE_BATCH_PERSISTENT or NE_BATCH_WAVEFRONT."""

    def __init__(self, ctx: Context, jobs: list):
        self.ctx = ctx
        self.jobs = jobs
        arr = (Job * len(jobs))()
        for k, j in enumerate(jobs):
            a = arr[k]
            a.scene = j["scene"].h.value if isinstance(j["scene"].h, C.c_void_p) else j["scene"].h
            a.camera = j["camera"].desc()
            a.width, a.height, a.spp, a.max_bounces = j["width"], j["height"], j["spp"], j["max_bounces"]
            a.seed, a.job = j.get("seed", 0), j.get("job", 0)
            g = j.get("guide")
            a.guide = (g.h.value if isinstance(g.h, C.c_void_p) else g.h) if g is not None else None
        self._keep = [j["scene"] for j in jobs] + [j["guide"] for j in jobs if j.get("guide") is not None]
        h = C.c_void_p()
        ctx.check(ctx._lib.ne_batch_create(ctx.h, arr, len(jobs), C.byref(h)))
        self.h = h

    def autotune(self, spp_divisor: int = 8):
        cfg, ms = BatchConfig(), C.c_float()
        self.ctx.check(self.ctx._lib.ne_batch_autotune(self.h, spp_divisor, C.byref(cfg), C.byref(ms)))
        return (cfg.tile_w, cfg.tile_h, cfg.samples_per_item, cfg.executor), float(ms.value)

    def run(self, cfg=None) -> float:
        ms = C.c_float()
        p = None
        if cfg is not None:
            c = BatchConfig(cfg[0], cfg[1], cfg[2], cfg[3] if len(cfg) > 3 else 0)
            p = C.byref(c)
        self.ctx.check(self.ctx._lib.ne_batch_run(self.h, p, C.byref(ms)))
        return float(ms.value)

    def job_times(self) -> np.ndarray:
        out = np.zeros(len(self.jobs), np.float32)
        self.ctx.check(self.ctx._lib.ne_batch_job_times(self.h, _fp(out)))
        return out

    def read(self, k: int) -> np.ndarray:
        j = self.jobs[k]
        out = np.zeros((j["height"], j["width"], 3), np.float64)
        self.ctx.check(self.ctx._lib.ne_batch_read(self.h, k, _dp(out)))
        return out

    def close(self):
        if self.h:
            self.ctx._lib.ne_batch_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Guide:
    """The learned directional guide (ne_guide)."""

    def __init__(self, ctx: Context, log2_table: int = 19, base_res: int = 16, max_res: int = 512,
                 batch: int = 65536, lr_table: float = 1e-2, lr_mlp: float = 1e-3, beta1: float = 0.9,
                 beta2: float = 0.99, eps: float = 1e-8, seed: int = 0, guided_vertices: int = 1,
                 guide_prob: float = 0.25, cache_res: int = 0, cache_dirs: int = 4):
        self.ctx = ctx
        cfg = GuideConfig(log2_table, base_res, max_res, batch, lr_table, lr_mlp, beta1, beta2, eps, guided_vertices,
                          guide_prob, seed, cache_res, cache_dirs)
        h = C.c_void_p()
        ctx.check(ctx._lib.ne_guide_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        tn, mn = C.c_uint64(), C.c_int32()
        res = (C.c_int32 * 16)()
        ctx.check(ctx._lib.ne_guide_sizes(self.h, C.byref(tn), C.byref(mn), res))
        self.table_floats, self.mlp_floats, self.level_res = int(tn.value), int(mn.value), list(res)
        self.log2_table = log2_table

    def eval(self, queries: np.ndarray, raw: bool = False) -> np.ndarray:
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 6)
        out = np.zeros((len(q), 40), np.float32)
        self.ctx.check(self.ctx._lib.ne_guide_eval(self.h, len(q), _fp(q), 1 if raw else 0, _fp(out)))
        return out

    def train_step(self, samples: np.ndarray, update: bool = True) -> float:
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, 10)
        loss = C.c_float()
        self.ctx.check(self.ctx._lib.ne_guide_train_step(self.h, len(s), _fp(s), 1 if update else 0, C.byref(loss)))
        return float(loss.value)

    def params(self):
        t = np.zeros(self.table_floats, np.float32)
        m = np.zeros(self.mlp_floats, np.float32)
        self.ctx.check(self.ctx._lib.ne_guide_get_params(self.h, _fp(t), _fp(m)))
        return t, m

    def set_params(self, table=None, mlp=None):
        t = None if table is None else np.ascontiguousarray(table, np.float32)
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        self.ctx.check(self.ctx._lib.ne_guide_set_params(self.h, _fp(t) if t is not None else None,
                                                         _fp(m) if m is not None else None))

    def grads(self):
        t = np.zeros(self.table_floats, np.float32)
        m = np.zeros(self.mlp_floats, np.float32)
        self.ctx.check(self.ctx._lib.ne_guide_get_grads(self.h, _fp(t), _fp(m)))
        return t, m

    def cache(self) -> np.ndarray:
        """The tabulated proposal: (cache_res^3 * cache_dirs^2, 40) mixture records."""
        n = C.c_uint64()
        self.ctx.check(self.ctx._lib.ne_guide_get_cache(self.h, None, C.byref(n)))
        out = np.zeros((n.value, 40), np.float32)
        self.ctx.check(self.ctx._lib.ne_guide_get_cache(self.h, _fp(out), C.byref(n)))
        return out

    def stats(self) -> dict:
        s = GuideStats()
        self.ctx.check(self.ctx._lib.ne_guide_get_stats(self.h, C.byref(s)))
        return {"steps": int(s.steps), "samples": int(s.samples), "last_loss": float(s.last_loss)}

    def reset(self):
        self.ctx.check(self.ctx._lib.ne_guide_reset(self.h))

    def close(self):
        if self.h:
            self.ctx._lib.ne_guide_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
