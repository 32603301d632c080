"""Host-side scene builder: numpy arrays -> ne_scene_desc.

Follows what the reference's SceneReader builds (src/io/SceneReader.cpp:224-648):
  * instance transform = translate * eulerAngleXYZ(rotation in degrees) * scale
    (getTransform, src/utils/Math.h:848-859; eulerAngleXYZ = Rx * Ry * Rz);
  * rectangle: unit quad [-.5,.5]^2 at z = 0, OCS normal (0,0,-1)
    (SceneReader.cpp:407-419), normal taken to WCS as transform * (n, 0),
    normalised (src/primitives/InstancedModel.cpp:31-34);
  * sphere: centre at the instance position, radius as given (translation only,
    SceneReader.cpp:385);
  * volume: GridMedia on the unit cube [-.5,.5]^3 under the instance transform,
    or HomogeneousMedia (constant density) on the same cube.
The flattened, world-space arrays are the ABI input shared by the GPU library
and the test oracle.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from .abi import CameraDesc, SceneDesc, NE_MAT_GLOSSY, NE_MAT_EMITTER, NE_ENV_NONE, NE_ENV_MAP


def rot_xyz(deg) -> np.ndarray:
    rx, ry, rz = (math.radians(float(a)) for a in deg)
    cx, sx, cy, sy, cz, sz = math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry), math.cos(rz), math.sin(rz)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def transform(position=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1)) -> np.ndarray:
    M = np.eye(4)
    M[:3, :3] = rot_xyz(rotation) @ np.diag(np.asarray(scale, dtype=np.float64))
    M[:3, 3] = np.asarray(position, dtype=np.float64)
    return M


def _ptr(a: np.ndarray | None, ctype):
    if a is None or a.size == 0:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Camera:
    position: tuple = (0.0, 0.0, -2.0)
    look_at: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    vfov: float = 45.0
    aspect: float = 1.0

    def desc(self) -> CameraDesc:
        d = CameraDesc()
        d.position[:] = [float(x) for x in self.position]
        d.look_at[:] = [float(x) for x in self.look_at]
        d.up[:] = [float(x) for x in self.up]
        d.vfov = float(self.vfov)
        d.aspect = float(self.aspect)
        return d


@dataclass
class Scene:
    tri_v: list = field(default_factory=list)      # arrays (n, 9) float32
    tri_mat: list = field(default_factory=list)
    sph: list = field(default_factory=list)        # (cx, cy, cz, r)
    sph_mat: list = field(default_factory=list)
    rect: list = field(default_factory=list)       # 12 floats
    rect_mat: list = field(default_factory=list)
    mats: list = field(default_factory=list)       # 12 floats
    points: list = field(default_factory=list)     # point-light positions
    point_mat: list = field(default_factory=list)
    volume: dict | None = None
    sun: dict | None = None
    env: dict | None = None

    # ---- materials (SceneReader::processMaterial, src/io/SceneReader.cpp:67-222)
    def add_glossy(self, albedo, roughness, metallic=0.0) -> int:
        m = np.zeros(12, np.float32)
        m[0] = NE_MAT_GLOSSY
        m[1:4] = albedo
        m[4] = roughness
        m[5] = metallic
        self.mats.append(m)
        return len(self.mats) - 1

    def add_emitter(self, radiance) -> int:
        m = np.zeros(12, np.float32)
        m[0] = NE_MAT_EMITTER
        m[6:9] = radiance
        self.mats.append(m)
        return len(self.mats) - 1

    # ---- primitives (SceneReader::processPrimitives)
    def add_mesh(self, verts: np.ndarray, tris: np.ndarray, mat: int, xf: np.ndarray | None = None):
        v = np.asarray(verts, np.float64).reshape(-1, 3)
        if xf is not None:
            v = v @ xf[:3, :3].T + xf[:3, 3]
        t = np.asarray(tris, np.int64).reshape(-1, 3)
        tv = v[t].reshape(-1, 9).astype(np.float32)
        self.tri_v.append(tv)
        self.tri_mat.append(np.full(len(tv), mat, np.int32))

    def add_sphere(self, center, radius, mat: int):
        self.sph.append(np.array([*center, radius], np.float32))
        self.sph_mat.append(mat)

    def add_rectangle(self, mat: int, position=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1)):
        M = transform(position, rotation, scale)
        p0 = M @ np.array([-0.5, -0.5, 0.0, 1.0])
        eu = M[:3, :3] @ np.array([1.0, 0.0, 0.0])
        ev = M[:3, :3] @ np.array([0.0, 1.0, 0.0])
        n = M[:3, :3] @ np.array([0.0, 0.0, -1.0])
        n = n / np.linalg.norm(n)
        self.rect.append(np.concatenate([p0[:3], eu, ev, n]).astype(np.float32))
        self.rect_mat.append(mat)

    def add_point_light(self, position, mat: int):
        """A point primitive with an emitter material (Point, src/primitives/Point.cpp):
        a delta light of intensity = the material's radiance."""
        self.points.append(np.array(position, np.float32))
        self.point_mat.append(mat)

    def set_volume(self, grid: np.ndarray, density_mult: float, sigma_s, sigma_a, g: float,
                   position=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1), maj_res: int = 16,
                   empty_skip: bool = True):
        """empty_skip: the majorant DDA steps over runs of empty cells without
        reading them (same image; off only to measure what it saves)."""
        grid = np.ascontiguousarray(grid, dtype=np.float32)
        assert grid.ndim == 3, "grid must be (D, H, W), x fastest"
        M = transform(position, rotation, scale)
        self.volume = dict(grid=grid, mult=float(density_mult), sigma_s=tuple(sigma_s), sigma_a=tuple(sigma_a),
                           g=float(g), xf=M, inv=np.linalg.inv(M), maj_res=int(maj_res),
                           empty_skip=bool(empty_skip))

    def set_homogeneous_volume(self, density: float, sigma_s, sigma_a, g: float,
                               position=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1)):
        """HomogeneousMedia (src/materials/HomogeneousMedia.cpp:5-51) on the unit
        cube under the instance transform; the material's "density" is the
        constant density."""
        M = transform(position, rotation, scale)
        self.volume = dict(grid=None, mult=float(density), sigma_s=tuple(sigma_s), sigma_a=tuple(sigma_a),
                           g=float(g), xf=M, inv=np.linalg.inv(M), maj_res=1)

    def set_sun(self, direction, radiance=(1.0, 1.0, 1.0)):
        d = np.asarray(direction, np.float64)
        self.sun = dict(dir=d / np.linalg.norm(d), E=tuple(radiance))

    def set_env(self, kind: int, c0=(0, 0, 0), c1=(0, 0, 0), c2=(0, 0, 0)):
        self.env = dict(kind=int(kind), c0=tuple(c0), c1=tuple(c1), c2=tuple(c2))

    def set_env_map(self, rgb: np.ndarray):
        """InfiniteAreaLight (src/lights/InfiniteAreaLight.h:8-130): an (H, W, 3)
        latitude-longitude radiance map, row 0 at theta = 0 of its z-up frame."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        assert rgb.ndim == 3 and rgb.shape[2] == 3, "env map must be (H, W, 3)"
        self.env = dict(kind=NE_ENV_MAP, c0=(0, 0, 0), c1=(0, 0, 0), c2=(0, 0, 0), map=rgb)

    # ---- flattening
    def arrays(self) -> dict:
        cat = lambda xs, w, dt: (np.ascontiguousarray(np.concatenate(xs).reshape(-1, w), dt) if xs
                                 else np.zeros((0, w), dt))
        a = dict(
            tri_v=cat(self.tri_v, 9, np.float32), tri_mat=(np.ascontiguousarray(np.concatenate(self.tri_mat), np.int32)
                                                        if self.tri_mat else np.zeros(0, np.int32)),
            sph=cat(self.sph, 4, np.float32), sph_mat=np.asarray(self.sph_mat, np.int32),
            rect=cat(self.rect, 12, np.float32), rect_mat=np.asarray(self.rect_mat, np.int32),
            mat=cat(self.mats, 12, np.float32),
            point=cat(self.points, 3, np.float32), point_mat=np.asarray(self.point_mat, np.int32),
        )
        if self.volume is not None and self.volume["grid"] is not None:
            a["grid"] = self.volume["grid"]
        if self.env is not None and "map" in self.env:
            a["env_map"] = self.env["map"]
        return a

    def desc(self) -> tuple[SceneDesc, dict]:
        """Returns (ne_scene_desc, keep-alive dict of the arrays it points into)."""
        a = self.arrays()
        d = SceneDesc()
        d.tri_v, d.tri_mat, d.n_tri = _ptr(a["tri_v"], C.c_float), _ptr(a["tri_mat"], C.c_int32), len(a["tri_v"])
        d.sph, d.sph_mat, d.n_sph = _ptr(a["sph"], C.c_float), _ptr(a["sph_mat"], C.c_int32), len(a["sph"])
        d.rect, d.rect_mat, d.n_rect = _ptr(a["rect"], C.c_float), _ptr(a["rect_mat"], C.c_int32), len(a["rect"])
        d.mat, d.n_mat = _ptr(a["mat"], C.c_float), len(a["mat"])
        d.point, d.point_mat, d.n_point = _ptr(a["point"], C.c_float), _ptr(a["point_mat"], C.c_int32), len(a["point"])
        if self.volume is not None:
            v = self.volume
            d.has_volume = 1
            d.vol_xf[:] = [float(x) for x in v["xf"].reshape(-1)]
            d.vol_inv[:] = [float(x) for x in v["inv"].reshape(-1)]
            if v["grid"] is None:
                d.vol_homogeneous = 1
            else:
                D, H, W = v["grid"].shape
                d.vol_res[:] = [W, H, D]
                d.grid = _ptr(a["grid"], C.c_float)
            d.vol_mult = v["mult"]
            d.vol_sigma_s[:] = v["sigma_s"]
            d.vol_sigma_a[:] = v["sigma_a"]
            d.vol_g = v["g"]
            d.maj_res = v["maj_res"]
            d.maj_no_skip = 0 if v.get("empty_skip", True) else 1
        if self.sun is not None:
            d.has_sun = 1
            d.sun_dir[:] = [float(x) for x in self.sun["dir"]]
            d.sun_E[:] = self.sun["E"]
        if self.env is not None:
            d.env_type = self.env["kind"]
            d.env_c0[:] = self.env["c0"]
            d.env_c1[:] = self.env["c1"]
            d.env_c2[:] = self.env["c2"]
            if "map" in self.env:
                d.env_map = _ptr(a["env_map"], C.c_float)
                d.env_h, d.env_w = self.env["map"].shape[:2]
        else:
            d.env_type = NE_ENV_NONE
        return d, a
