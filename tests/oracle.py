"""ctypes view of the oracle (oracle/liboracle.so, the CPU restatement of the
reference path tracer).  Test infrastructure: used only as the checker."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
_F, _I, _D, _U = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_uint32)
_L = None


def build():
    res = subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)


def lib():
    global _L
    if _L is None:
        p = ORACLE_DIR / "liboracle.so"
        src = [ORACLE_DIR / "ne_oracle.hpp", ORACLE_DIR / "oracle_capi.cpp"]
        if not p.exists() or any(s.stat().st_mtime > p.stat().st_mtime for s in src):
            build()
        L = C.CDLL(str(p))
        sig = {
            "neo_sizeof_scene_desc": (C.c_int, []),
            "neo_sizeof_camera_desc": (C.c_int, []),
            "neo_philox": (None, [_U, _U, _U]),
            "neo_rng_uniforms": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _F]),
            "neo_aabb_isect": (None, [C.c_int, _F, _F, _I, _F, _F, _F]),
            "neo_tri_isect": (None, [C.c_int, _F, _F, _I, _F, _F]),
            "neo_sphere_isect": (None, [C.c_int, _F, _F, _I, _F, _F]),
            "neo_rect_isect": (None, [C.c_int, _F, _F, _I, _F]),
            "neo_triangle_area": (None, [C.c_int, _F, _F]),
            "neo_barycentric": (None, [C.c_int, _F, _F, _F]),
            "neo_triangle_contains": (None, [C.c_int, _F, _F, _I]),
            "neo_onb": (None, [C.c_int, _F, _F, _F]),
            "neo_to_lcs": (None, [C.c_int, _F, _F, _F, _F, _F]),
            "neo_to_world": (None, [C.c_int, _F, _F, _F, _F, _F]),
            "neo_sample_unit_sphere": (None, [C.c_int, _F, _F]),
            "neo_hg_eval": (None, [C.c_int, C.c_float, _F, _F, _F]),
            "neo_hg_sample_local": (None, [C.c_int, C.c_float, _F, _F]),
            "neo_ggx_D": (None, [C.c_int, _F, _F, _F]),
            "neo_ggx_G": (None, [C.c_int, _F, _F, _F, _F]),
            "neo_ggx_pdf": (None, [C.c_int, _F, _F, _F, _F]),
            "neo_ggx_sample_h": (None, [C.c_int, _F, _F, _F]),
            "neo_glossy_eval": (None, [C.c_int, _F, _F, _F, _F, _F, _F]),
            "neo_interp_density": (None, [C.c_int, C.c_int, C.c_int, _F, C.c_int, _F, _F]),
            "neo_max_density": (C.c_float, [C.c_int, C.c_int, C.c_int, _F]),
            "neo_delta_track": (None, [C.c_int, C.c_int, C.c_int, _F, C.c_float, C.c_float, C.c_int, _F, _F,
                                       C.c_uint64, _I, _F]),
            "neo_ratio_track": (None, [C.c_int, C.c_int, C.c_int, _F, C.c_float, C.c_float, C.c_int, _F, _F,
                                       C.c_uint64, _F]),
            "neo_post_processing": (None, [C.c_int, _F, _F]),
            "neo_homogeneous_tr": (None, [C.c_int, _F, _F, _F, _F]),
            "neo_world_env_sample": (None, [C.c_void_p, C.c_int, _F, _F, _F, _F]),
            "neo_world_env_eval": (None, [C.c_void_p, C.c_int, _F, _F, _F]),
            "neo_ref_dist2d": (None, [_F, C.c_int, C.c_int, C.c_int, C.c_int, _F, _F, _F]),
            "neo_homogeneous_sample": (None, [C.c_int, _F, _F, C.c_float, _F, C.c_uint64, _I, _F, _F]),
            "neo_camera_dirs": (None, [C.c_void_p, C.c_int, _F, _F]),
            "neo_tile_rect": (None, [C.c_int, C.c_int, C.c_int, _I]),
            "neo_hg_p": (None, [C.c_int, _F, _F, _F]),
            "neo_hg_sample": (None, [C.c_int, _F, _F, _F, _F]),
            "neo_glossy_c": (None, [C.c_int, _F, _F, _F, _F, _F, _F, _F]),
            "neo_world_create": (C.c_void_p, [C.c_void_p, C.c_int]),
            "neo_world_destroy": (None, [C.c_void_p]),
            "neo_world_bvh_nodes": (C.c_int, [C.c_void_p]),
            "neo_world_majorants": (None, [C.c_void_p, _F]),
            "neo_world_closest": (None, [C.c_void_p, C.c_int, _F, C.c_float, C.c_float, _F, _I, _I]),
            "neo_world_closest_brute": (None, [C.c_void_p, C.c_int, _F, C.c_float, C.c_float, _F, _I, _I]),
            "neo_world_transmittance": (None, [C.c_void_p, C.c_int, _F, _F, C.c_uint64, _D]),
            "neo_world_delta_track": (None, [C.c_void_p, C.c_int, _F, C.c_uint64, _I, _D]),
            "neo_li": (None, [C.c_void_p, C.c_int, _F, C.c_int, C.c_uint64, C.c_uint32, _D]),
            "neo_render_pixels": (None, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, _I, C.c_int, C.c_int,
                                         C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, _D, _D]),
            "neo_render": (None, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, _D,
                                  C.POINTER(C.c_uint64)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _L = L
    return _L


def f(a):
    return np.ascontiguousarray(a, np.float32)


def fp(a):
    return a.ctypes.data_as(_F)


def ip(a):
    return a.ctypes.data_as(_I)


def dp(a):
    return a.ctypes.data_as(_D)


# ---- thin numpy wrappers -------------------------------------------------------
def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().neo_philox(c.ctypes.data_as(_U), k.ctypes.data_as(_U), o.ctypes.data_as(_U))
    return o


def rng_uniforms(seed, pixel, sample, job, n):
    o = np.zeros(n, np.float32)
    lib().neo_rng_uniforms(seed, pixel, sample, job, n, fp(o))
    return o


def aabb(rays, boxes):
    rays, boxes = f(rays).reshape(-1, 6), f(boxes).reshape(-1, 6)
    n = len(rays)
    hit, tn, tf, pt = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
    lib().neo_aabb_isect(n, fp(rays), fp(boxes), ip(hit), fp(tn), fp(tf), fp(pt))
    return hit, tn, tf, pt


def triangle(rays, tris):
    rays, tris = f(rays).reshape(-1, 6), f(tris).reshape(-1, 9)
    n = len(rays)
    hit, t, nrm = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
    lib().neo_tri_isect(n, fp(rays), fp(tris), ip(hit), fp(t), fp(nrm))
    return hit, t, nrm


def sphere(rays, sph):
    rays, sph = f(rays).reshape(-1, 6), f(sph).reshape(-1, 4)
    n = len(rays)
    hit, tn, tf = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    lib().neo_sphere_isect(n, fp(rays), fp(sph), ip(hit), fp(tn), fp(tf))
    return hit, tn, tf


def rectangle(rays, rects):
    rays, rects = f(rays).reshape(-1, 6), f(rects).reshape(-1, 9)
    n = len(rays)
    hit, t = np.zeros(n, np.int32), np.zeros(n, np.float32)
    lib().neo_rect_isect(n, fp(rays), fp(rects), ip(hit), fp(t))
    return hit, t


def vec_op(name, n, *arrays, out_shape):
    arrs = [f(a) for a in arrays]
    out = np.zeros(out_shape, np.float32)
    getattr(lib(), name)(n, *[fp(a) for a in arrs], fp(out))
    return out


def homogeneous_tr(ext, density, dist):
    ext, density, dist = f(ext).reshape(-1, 3), f(density), f(dist)
    out = np.zeros((len(ext), 3), np.float32)
    lib().neo_homogeneous_tr(len(ext), fp(ext), fp(density), fp(dist), fp(out))
    return out


def homogeneous_sample(sig_s, sig_a, density, tint, seed=0):
    tint = f(tint).reshape(-1, 2)
    n = len(tint)
    sampled, t, w = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
    lib().neo_homogeneous_sample(n, fp(f(sig_s)), fp(f(sig_a)), float(density), fp(tint), seed, ip(sampled), fp(t),
                                 fp(w))
    return sampled, t, w


def ref_dist2d(func, cond_n, u2):
    func = f(func)
    H, W = func.shape
    u2 = f(u2).reshape(-1, 2)
    uv, pdf = np.zeros((len(u2), 2), np.float32), np.zeros(len(u2), np.float32)
    lib().neo_ref_dist2d(fp(func), W, H, cond_n, len(u2), fp(u2), fp(uv), fp(pdf))
    return uv, pdf


class World:
    """Oracle scene (BVH, volume, lights) built from the same ne_scene_desc."""

    def __init__(self, scene, fp64: bool = False):
        self.desc, self.keep = scene.desc()
        assert lib().neo_sizeof_scene_desc() == C.sizeof(self.desc)
        self.h = lib().neo_world_create(C.byref(self.desc), 1 if fp64 else 0)

    def __del__(self):
        if getattr(self, "h", None):
            lib().neo_world_destroy(self.h)
            self.h = None

    def closest(self, rays, tmin=1e-4, tmax=np.inf, brute=False):
        rays = f(rays).reshape(-1, 6)
        n = len(rays)
        t, ty, ix = np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        fn = lib().neo_world_closest_brute if brute else lib().neo_world_closest
        fn(self.h, n, fp(rays), tmin, tmax, fp(t), ip(ty), ip(ix))
        return t, ty, ix

    def transmittance(self, rays, tmax, seed=0):
        rays = f(rays).reshape(-1, 6)
        tmax = f(tmax)
        out = np.zeros(len(rays), np.float64)
        lib().neo_world_transmittance(self.h, len(rays), fp(rays), fp(tmax), seed, dp(out))
        return out

    def delta_track(self, rays, seed=0):
        rays = f(rays).reshape(-1, 6)
        sc, t = np.zeros(len(rays), np.int32), np.zeros(len(rays), np.float64)
        lib().neo_world_delta_track(self.h, len(rays), fp(rays), seed, ip(sc), dp(t))
        return sc, t

    def li(self, rays, max_bounces, seed=0, sample=0):
        """Radiance along each ray (n, 6) -> (n, 3) float64; ray i uses stream (seed; i, sample, 0)."""
        rays = f(rays).reshape(-1, 6)
        out = np.zeros((len(rays), 3), np.float64)
        lib().neo_li(self.h, len(rays), fp(rays), max_bounces, seed, sample, dp(out))
        return out

    def render_pixels(self, camera, W, H, xy, spp, max_bounces, seed=0, sample_offset=0, job=0, threads=8):
        """Per-pixel FP64 (sum, sum of squares) of spp camera samples at pixels xy (n, 2),
        with the streams and jitter of render()."""
        xy = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
        n = len(xy)
        s1, s2 = np.zeros((n, 3), np.float64), np.zeros((n, 3), np.float64)
        cd = camera.desc()
        lib().neo_render_pixels(self.h, C.byref(cd), W, H, n, ip(xy), spp, max_bounces, seed, sample_offset, job,
                                threads, dp(s1), dp(s2))
        return s1, s2

    def env_sample(self, u2):
        u2 = f(u2).reshape(-1, 2)
        n = len(u2)
        d, p, le = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
        lib().neo_world_env_sample(self.h, n, fp(u2), fp(d), fp(p), fp(le))
        return d, p, le

    def env_eval(self, dirs):
        dirs = f(dirs).reshape(-1, 3)
        n = len(dirs)
        p, le = np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
        lib().neo_world_env_eval(self.h, n, fp(dirs), fp(p), fp(le))
        return p, le

    def majorants(self, mres):
        out = np.zeros(mres ** 3, np.float32)
        lib().neo_world_majorants(self.h, fp(out))
        return out

    def render(self, camera, W, H, spp, max_bounces, rect=None, seed=0, sample_offset=0, job=0, threads=8,
               stats=False):
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, W, H)
        s = np.zeros((H, W, 3), np.float64)
        cd = camera.desc()
        st = np.zeros(2, np.uint64)
        lib().neo_render(self.h, C.byref(cd), W, H, x0, y0, x1, y1, spp, max_bounces, seed, sample_offset, job,
                         threads, dp(s), st.ctypes.data_as(C.POINTER(C.c_uint64)) if stats else None)
        return (s, st) if stats else s
