"""The reference's renderer interfaces over the GPU backend.

  * OfflineEngine  -- src/core/OfflineEngine.{h,cpp}: constructed from
    (Camera, SceneSettings, Scene); renderTile(camera, index, finished) renders
    tile `index` of the 40 x 10 grid and writes post-processed RGB into
    `pixels` (row-major, x fastest), then sets finished; coreLoop renders every
    tile and writes a 16-bit PPM like the reference's headless loop
    (OfflineEngine.cpp:78-145).
  * GpuPathIntegrator -- src/integrators/Integrator.h:13-25: Li(ray) returns the
    radiance of one path; a batch form takes many rays at once.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from dataclasses import dataclass

import numpy as np

from .abi import NE_FP32, NE_IMAGE_PPM16, check, lib
from .runtime import Context


@dataclass
class SceneSettings:
    """src/core/Settings.h:9-14"""
    resolution: tuple = (512, 512)
    spp: int = 1
    bounces: int = 1
    hdr: bool = False


class OfflineEngine:
    number_of_tiles = (40, 10)  # OfflineEngine.cpp:19

    def __init__(self, camera, settings: SceneSettings, scene, ctx: Context | None = None, seed: int = 0):
        self.ctx = ctx or Context(0)
        self.camera, self.settings, self.scene = camera, settings, scene
        self.gscene = self.ctx.upload(scene)
        W, H = settings.resolution
        self.tile_size = (W // 40, H // 10)  # integer division, OfflineEngine.cpp:20-21
        self.film = self.ctx.film(W, H)
        self.pixels = np.zeros((H, W, 3), np.float32)
        self.seed = seed
        self._lock = threading.Lock()

    def renderTile(self, cam, index: int, finished: threading.Event | None = None):  # noqa: N802 (reference name)
        from .abi import RenderParams
        import ctypes as C
        W, H = self.settings.resolution
        p = RenderParams()
        p.width, p.height = W, H
        p.spp, p.max_bounces, p.seed, p.precision = self.settings.spp, self.settings.bounces, self.seed, NE_FP32
        cd = cam.desc()
        with self._lock:
            flat = self.pixels.reshape(-1)
            self.ctx.check(self.ctx._lib.ne_render_tile(self.ctx.h, self.gscene.h, C.byref(cd), C.byref(p), index,
                                                        self.film.h, flat.ctypes.data_as(C.POINTER(C.c_float))))
        if finished is not None:
            finished.set()

    def coreLoop(self, path: str = "output.ppm"):  # noqa: N802 (reference name)
        for i in range(self.number_of_tiles[0] * self.number_of_tiles[1]):
            self.renderTile(self.camera, i)
        write_ppm16(path, self.pixels)


def write_image(path, pixels: np.ndarray, fmt: int = NE_IMAGE_PPM16) -> None:
    """Writes an (H, W, 3) float image with the library's writers
    (ne_image_write): NE_IMAGE_PPM16 is OfflineEngine::coreLoop's output.ppm
    (OfflineEngine.cpp:80-137), NE_IMAGE_PNG / NE_IMAGE_EXR follow saveImage
    (src/materials/Texture.h:44-69), NE_IMAGE_PFM keeps linear floats."""
    img = np.ascontiguousarray(pixels, np.float32)
    H, W, _ = img.shape
    check(lib().ne_image_write(os.fsencode(str(path)), int(fmt), W, H, img.ctypes.data_as(C.POINTER(C.c_float))))


def write_ppm16(path, pixels: np.ndarray) -> None:
    """16-bit binary PPM as OfflineEngine::coreLoop writes it."""
    write_image(path, pixels, NE_IMAGE_PPM16)


class GpuPathIntegrator:
    """Integrator::Li over the GPU (one path per ray, FP32)."""

    def __init__(self, scene, bounces: int, ctx: Context | None = None, seed: int = 0):
        self.ctx = ctx or Context(0)
        self.gscene = self.ctx.upload(scene)
        self.bounces, self.seed = bounces, seed
        self._sample = 0

    def Li(self, origin, direction) -> np.ndarray:  # noqa: N802 (reference name)
        r = np.asarray([*origin, *direction], np.float32).reshape(1, 6)
        return self.Li_batch(r)[0]

    def Li_batch(self, rays: np.ndarray) -> np.ndarray:  # noqa: N802
        out = self.ctx.li(self.gscene, rays, self.bounces, self.seed, self._sample)
        self._sample += 1
        return out

    def clone(self) -> "GpuPathIntegrator":
        other = object.__new__(GpuPathIntegrator)
        other.ctx, other.gscene, other.bounces, other.seed, other._sample = (
            self.ctx, self.gscene, self.bounces, self.seed, self._sample)
        return other
