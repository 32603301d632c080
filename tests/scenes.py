"""Small seeded scenes shared by the parity tests (sizes the oracle finishes
in seconds)."""
from __future__ import annotations

import numpy as np

from narvalengine_amd import configs
from narvalengine_amd.abi import NE_ENV_CONSTANT, NE_ENV_SKY
from narvalengine_amd.scene import Camera, Scene


def small_volume(res=32, seed=0, maj_res=4, env=True, sun=True, mult=20.0, albedo=0.99, g=0.85) -> Scene:
    s = Scene()
    s.set_volume(configs.fbm_grid(res, seed), mult, (albedo,) * 3, (1 - albedo,) * 3, g, maj_res=maj_res)
    if sun:
        s.set_sun(-np.array([1.0, 2.0, -0.7]), (1.0, 1.0, 1.0))
    if env:
        s.set_env(NE_ENV_CONSTANT, c0=(0.2, 0.25, 0.35))
    return s


def homogeneous_box(density=3.0, albedo=0.9, g=0.6, rotation=(20, 35, 0)) -> Scene:
    """A rotated, non-uniformly scaled homogeneous medium over a floor, lit by
    an area light, the sun and a sky."""
    s = Scene()
    s.set_homogeneous_volume(density, (albedo,) * 3, (1 - albedo,) * 3, g, position=(0, 0.1, 0),
                             rotation=rotation, scale=(0.8, 0.5, 0.6))
    floor = s.add_glossy((0.6, 0.6, 0.6), 0.7)
    s.add_rectangle(floor, position=(0, -0.5, 0), rotation=(90, 0, 0), scale=(4, 4, 1))
    light = s.add_emitter((8.0, 8.0, 8.0))
    s.add_rectangle(light, position=(0.3, 1.2, -0.2), rotation=(-90, 0, 0), scale=(0.5, 0.5, 1))
    s.set_sun(-np.array([1.0, 2.0, -0.7]), (1.0, 1.0, 1.0))
    s.set_env(NE_ENV_SKY, c0=(0.2, 0.3, 0.6), c1=(0.6, 0.6, 0.6), c2=(0.1, 0.1, 0.1))
    return s


def emitter_mix() -> Scene:
    """Every emitter shape: a rectangle, two spheres (one small and far, one
    close to the walls), an emissive triangle pair and a point light, over
    glossy walls."""
    s = Scene()
    wall = s.add_glossy((0.7, 0.7, 0.7), 0.6)
    red = s.add_glossy((0.8, 0.2, 0.2), 0.3, 0.5)
    s.add_rectangle(wall, position=(0, 0, 0), rotation=(90, 0, 0), scale=(2, 2, 1))
    s.add_rectangle(wall, position=(0, 1, 1), rotation=(0, 0, 0), scale=(2, 2, 1))
    s.add_rectangle(red, position=(-1, 1, 0), rotation=(0, 90, 0), scale=(2, 2, 1))
    s.add_rectangle(s.add_emitter((4.0, 4.0, 4.0)), position=(0, 1.98, 0), rotation=(-90, 0, 0),
                    scale=(0.3, 0.3, 1))
    s.add_sphere((0.5, 0.4, 0.2), 0.2, s.add_emitter((6.0, 3.0, 1.0)))
    s.add_sphere((-0.6, 1.5, 0.6), 0.05, s.add_emitter((60.0, 60.0, 80.0)))
    s.add_sphere((-0.3, 0.3, -0.2), 0.3, red)
    q = np.array([[0.7, 0.9, 0.99], [0.95, 0.9, 0.99], [0.95, 1.3, 0.99], [0.7, 1.3, 0.99]])
    s.add_mesh(q, np.array([[0, 1, 2], [0, 2, 3]]), s.add_emitter((2.0, 5.0, 2.0)))
    s.add_point_light((0.2, 1.6, -0.3), s.add_emitter((0.8, 0.8, 0.6)))
    return s


def env_map_scene(res=24, seed=0) -> Scene:
    """A volume over a glossy floor under an importance-sampled environment
    map with a bright patch (InfiniteAreaLight) and the sun."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.set_volume(configs.fbm_grid(res, seed), 12.0, (0.95,) * 3, (0.05,) * 3, 0.7, maj_res=4)
    floor = s.add_glossy((0.6, 0.55, 0.5), 0.5)
    s.add_rectangle(floor, position=(0, -0.5, 0), rotation=(90, 0, 0), scale=(6, 6, 1))
    m = rng.gamma(0.8, 0.3, (32, 64, 3)).astype(np.float32)
    m[6:9, 20:26] *= 60.0
    s.set_env_map(m)
    s.set_sun(-np.array([1.0, 2.0, -0.7]), (0.5, 0.5, 0.5))
    return s


def small_cloud(res=48, maj_res=6) -> Scene:
    return configs.cloud_scene(configs.cumulus_grid(res, 2), maj_res=maj_res)


def small_surface(subdiv=3, seed=1) -> Scene:
    rng = np.random.default_rng(seed)
    s = Scene()
    mats = [s.add_glossy(rng.uniform(0.1, 0.9, 3), float(rng.uniform(0.05, 0.8))) for _ in range(9)]
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
    v, t = configs.icosphere(subdiv, seed, 0.1)
    xf = np.eye(4)
    xf[:3, :3] *= 0.45
    xf[:3, 3] = (0.0, 0.55, 0.35)
    s.add_mesh(v, t, mats[8], xf)
    return s


BOX_CAMERA = Camera(position=(0.0, 1.0, -3.2), look_at=(0, 1, 0), up=(0, 1, 0), vfov=40.0, aspect=1.0)
VOLUME_CAMERA = Camera(position=(0.0, 0.25, -1.8), look_at=(0, 0, 0), up=(0, 1, 0), vfov=45.0, aspect=1.0)


def triangle_soup(n=300, seed=3) -> Scene:
    rng = np.random.default_rng(seed)
    s = Scene()
    m = s.add_glossy((0.5, 0.5, 0.5), 0.3)
    c = rng.uniform(-1, 1, (n, 1, 3))
    v = c + rng.normal(0, 0.15, (n, 3, 3))
    s.tri_v.append(v.reshape(-1, 9).astype(np.float32))
    s.tri_mat.append(np.zeros(n, np.int32))
    for k in range(6):
        s.add_sphere(tuple(rng.uniform(-1, 1, 3)), float(rng.uniform(0.05, 0.3)), m)
    for k in range(5):
        s.add_rectangle(m, position=tuple(rng.uniform(-1, 1, 3)), rotation=tuple(rng.uniform(0, 180, 3)),
                        scale=(float(rng.uniform(0.2, 0.8)), float(rng.uniform(0.2, 0.8)), 1))
    return s


def random_rays(n, seed=4, origin_box=2.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-origin_box, origin_box, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def rays_towards(n, target_box=0.5, origin_radius=2.5, seed=5):
    """Rays from a sphere of origins aimed at random points of a box."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * origin_radius
    tgt = rng.uniform(-target_box, target_box, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)
