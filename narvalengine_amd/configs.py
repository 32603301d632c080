"""The benchmark workloads as seeded synthetic scenes.

  volume128   128^3 FP32 4-octave Perlin fBm cloud (seed 0), sigma_t = 20 d,
              albedo 0.99, HG g = 0.85, unit sun + constant sky, pinhole
              512x512, 64 spp, 16 bounces, delta tracking.
  surface1m   box of 5 rectangles + 3 spheres + 1,310,720-triangle icosphere
              (8 subdivisions, vertices displaced by 0.1 fBm, seed 1), GGX
              roughness U[0.05, 0.8], albedo U[0.1, 0.9], one rectangle
              emitter of radiance 15, 1024x1024, 256 spp, 8 bounces, SAH BVH.
  cloud512    512^3 FP32 cumulus (Worley - fBm in an ellipsoid, seed 2), 16^3
              majorant cells, albedo 0.999, g = 0.877, sun + sky gradient,
              1920x1080, 256 spp, 128 bounces, ratio-tracked shadow rays.
  batch48     48 jobs (grids 64^3/128^3/256^3 built like volume128 from seeds
              1..48; resolutions 256^2 / 512^2 / 1280x720; spp 16/64/256).
"""
from __future__ import annotations

import ctypes as C
import functools
from dataclasses import dataclass

import numpy as np

from .abi import lib, check, NE_ENV_CONSTANT, NE_ENV_SKY
from .scene import Camera, Scene


@dataclass
class Workload:
    name: str
    scene: Scene
    camera: Camera
    width: int
    height: int
    spp: int
    max_bounces: int


@functools.lru_cache(maxsize=8)
def fbm_grid(res: int, seed: int, base_freq: float = 4.0) -> np.ndarray:
    g = np.zeros((res, res, res), np.float32)
    check(lib().ne_gen_fbm_grid(res, seed, base_freq, g.ctypes.data_as(C.POINTER(C.c_float))))
    g.setflags(write=False)
    return g


@functools.lru_cache(maxsize=2)
def cumulus_grid(res: int, seed: int) -> np.ndarray:
    g = np.zeros((res, res, res), np.float32)
    check(lib().ne_gen_cumulus_grid(res, seed, g.ctypes.data_as(C.POINTER(C.c_float))))
    g.setflags(write=False)
    return g


@functools.lru_cache(maxsize=2)
def icosphere(subdiv: int, seed: int, amp: float):
    nv, nt = C.c_uint32(), C.c_uint32()
    L = lib()
    null_f, null_u = C.cast(None, C.POINTER(C.c_float)), C.cast(None, C.POINTER(C.c_uint32))
    check(L.ne_gen_icosphere(subdiv, seed, amp, null_f, null_u, C.byref(nv), C.byref(nt)))
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    check(L.ne_gen_icosphere(subdiv, seed, amp, v.ctypes.data_as(C.POINTER(C.c_float)),
                             t.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(nv), C.byref(nt)))
    return v, t


def volume_scene(grid: np.ndarray, *, mult=20.0, albedo=0.99, g=0.85, maj_res=1) -> Scene:
    s = Scene()
    s.set_volume(grid, mult, (albedo,) * 3, (1.0 - albedo,) * 3, g, maj_res=maj_res)
    # DirectionalLight from `position` towards the origin (SceneReader.cpp:160-161)
    s.set_sun(direction=-np.array([1.0, 2.0, -0.7]), radiance=(1.0, 1.0, 1.0))
    s.set_env(NE_ENV_CONSTANT, c0=(0.2, 0.25, 0.35))
    return s


def volume128(res: int = 128, seed: int = 0, width=512, height=512, spp=64, bounces=16, maj_res=1) -> Workload:
    s = volume_scene(fbm_grid(res, seed), maj_res=maj_res)
    cam = Camera(position=(0.0, 0.25, -1.8), look_at=(0, 0, 0), up=(0, 1, 0), vfov=45.0, aspect=width / height)
    return Workload("volume128", s, cam, width, height, spp, bounces)


def surface1m(subdiv: int = 8, width=1024, height=1024, spp=256, bounces=8) -> Workload:
    rng = np.random.default_rng(1)
    s = Scene()
    mats = [s.add_glossy(rng.uniform(0.1, 0.9, 3), float(rng.uniform(0.05, 0.8))) for _ in range(9)]
    # box: floor, ceiling, back, left, right (2 x 2 x 2, open towards the camera)
    s.add_rectangle(mats[0], position=(0, 0, 0), rotation=(90, 0, 0), scale=(2, 2, 1))
    s.add_rectangle(mats[1], position=(0, 2, 0), rotation=(-90, 0, 0), scale=(2, 2, 1))
    s.add_rectangle(mats[2], position=(0, 1, 1), rotation=(0, 0, 0), scale=(2, 2, 1))
    s.add_rectangle(mats[3], position=(-1, 1, 0), rotation=(0, 90, 0), scale=(2, 2, 1))
    s.add_rectangle(mats[4], position=(1, 1, 0), rotation=(0, -90, 0), scale=(2, 2, 1))
    light = s.add_emitter((15.0, 15.0, 15.0))
    s.add_rectangle(light, position=(0, 1.98, 0.1), rotation=(-90, 0, 0), scale=(0.6, 0.6, 1))
    s.add_sphere((-0.6, 0.3, -0.3), 0.3, mats[5])
    s.add_sphere((0.62, 0.35, -0.1), 0.35, mats[6])
    s.add_sphere((-0.15, 0.2, -0.75), 0.2, mats[7])
    v, t = icosphere(subdiv, 1, 0.1)
    xf = np.eye(4)
    xf[:3, :3] *= 0.45
    xf[:3, 3] = (0.0, 0.55, 0.35)
    s.add_mesh(v, t, mats[8], xf)
    cam = Camera(position=(0.0, 1.0, -3.2), look_at=(0, 1, 0), up=(0, 1, 0), vfov=40.0, aspect=width / height)
    return Workload("surface1m", s, cam, width, height, spp, bounces)


def cloud_scene(grid: np.ndarray, maj_res=16) -> Scene:
    s = Scene()
    s.set_volume(grid, 40.0, (0.999,) * 3, (0.001,) * 3, 0.877, maj_res=maj_res)
    s.set_sun(direction=-np.array([0.6, 1.0, -0.4]), radiance=(2.6, 2.5, 2.3))
    s.set_env(NE_ENV_SKY, c0=(0.15, 0.25, 0.5), c1=(0.5, 0.55, 0.6), c2=(0.15, 0.14, 0.12))
    return s


def cloud512(res: int = 512, width=1920, height=1080, spp=256, bounces=128, maj_res=16) -> Workload:
    s = cloud_scene(cumulus_grid(res, 2), maj_res=maj_res)
    cam = Camera(position=(0.0, 0.05, -1.45), look_at=(0, 0.0, 0), up=(0, 1, 0), vfov=40.0, aspect=width / height)
    return Workload("cloud512", s, cam, width, height, spp, bounces)


def batch48_table(seed: int = 48):
    """(grid res, seed, width, height, spp) for the 48 uneven jobs."""
    rng = np.random.default_rng(seed)
    res_choices = [64, 128, 256]
    img_choices = [(256, 256), (512, 512), (1280, 720)]
    spp_choices = [16, 64, 256]
    jobs = []
    for j in range(48):
        r = res_choices[int(rng.integers(3))]
        w, h = img_choices[int(rng.integers(3))]
        spp = spp_choices[int(rng.integers(3))]
        jobs.append((r, j + 1, w, h, spp))
    return jobs


def batch48(seed: int = 48, bounces: int = 16) -> list:
    """The 48 uneven jobs as workloads: each its own fBm grid (built like
    volume128 from its seed), image size and spp; medium, lights and camera
    placement as volume128."""
    out = []
    for res, gseed, w, h, spp in batch48_table(seed):
        s = volume_scene(fbm_grid(res, gseed), maj_res=1)
        cam = Camera(position=(0.0, 0.25, -1.8), look_at=(0, 0, 0), up=(0, 1, 0), vfov=45.0, aspect=w / h)
        out.append(Workload(f"batch48/job{gseed}", s, cam, w, h, spp, bounces))
    return out


CONFIGS = {"volume128": volume128, "surface1m": surface1m, "cloud512": cloud512}
