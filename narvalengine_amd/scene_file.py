"""Scene files through the library's reader (ne_scene_file_*, ne_vol_read).

  * SceneFile   -- one JSON scene file (version 0.0.1a) read and flattened by
    libnarval_gpu.so: .desc() feeds Context.upload, .camera / .settings hold
    the main camera and SceneSettings, .warnings lists lenient substitutions.
  * SceneReader -- the reference's interface (src/io/SceneReader.h:8-40):
    SceneReader(filePath, absolutePath), loadScene, getScene, getMainCamera,
    getSettings.  A relative filePath, like every asset path inside the file,
    is taken relative to the resources directory (the reference's
    RESOURCES_DIR).
  * read_vol    -- the text .vol grid reader (ResourceManager::loadVolasTexture,
    src/core/ResourceManager.cpp:222-286) as a (D, H, W) float32 array.
  * write_vol / write_f32 / write_pfm / write_hdr -- writers of the grid and
    environment-image formats the reader takes.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import CameraDesc, LoadOptions, NarvalError, SceneDesc, SceneSettings as _Settings, check, lib
from .abi import NE_LOAD_LENIENT
from .engine import SceneSettings
from .scene import Camera


def _options(resources_dir, lenient, maj_res):
    o = LoadOptions()
    o.flags = NE_LOAD_LENIENT if lenient else 0
    o.maj_res = int(maj_res)
    o.resources_dir = os.fsencode(str(resources_dir)) if resources_dir is not None else None
    return o


class SceneFile:
    """A scene file read by the library.  Give `path` or `text` (JSON)."""

    def __init__(self, path=None, *, text: str | None = None, resources_dir=None, lenient: bool = False,
                 maj_res: int = 0):
        if (path is None) == (text is None):
            raise ValueError("give exactly one of path and text")
        self._lib = lib()
        opt = _options(resources_dir, lenient, maj_res)
        h = C.c_void_p()
        if path is not None:
            st = self._lib.ne_scene_file_read(os.fsencode(str(path)), C.byref(opt), C.byref(h))
        else:
            raw = text.encode("utf-8")
            st = self._lib.ne_scene_file_parse(raw, len(raw), C.byref(opt), C.byref(h))
        check(st)
        self.h = h
        self._desc = SceneDesc()
        cam = CameraDesc()
        st_ = _Settings()
        check(self._lib.ne_scene_file_get(self.h, C.byref(self._desc), C.byref(cam), C.byref(st_)))
        self.camera = Camera(position=tuple(cam.position), look_at=tuple(cam.look_at), up=tuple(cam.up),
                             vfov=cam.vfov, aspect=cam.aspect)
        self.settings = SceneSettings(resolution=(st_.width, st_.height), spp=st_.spp, bounces=st_.bounces,
                                      hdr=bool(st_.hdr))
        self.realtime = bool(st_.realtime)
        w = self._lib.ne_scene_file_warnings(self.h).decode()
        self.warnings = [x for x in w.split("\n") if x]

    def desc(self):
        """(ne_scene_desc, keep-alive): the arrays live in the library until close()."""
        if not self.h:
            raise NarvalError("scene file is closed")
        return self._desc, self

    def arrays(self) -> dict:
        """Copies of the flattened arrays (tri_v, tri_mat, sph, sph_mat, rect, rect_mat, mat, grid,
        env_map)."""
        d = self._desc

        def arr(ptr, n, w, dt):
            if n == 0 or not ptr:
                return np.zeros((0, w) if w > 1 else 0, dt)
            a = np.ctypeslib.as_array(ptr, shape=(n * w,)).astype(dt, copy=True)
            return a.reshape(n, w) if w > 1 else a

        out = dict(tri_v=arr(d.tri_v, d.n_tri, 9, np.float32), tri_mat=arr(d.tri_mat, d.n_tri, 1, np.int32),
                   sph=arr(d.sph, d.n_sph, 4, np.float32), sph_mat=arr(d.sph_mat, d.n_sph, 1, np.int32),
                   rect=arr(d.rect, d.n_rect, 12, np.float32), rect_mat=arr(d.rect_mat, d.n_rect, 1, np.int32),
                   mat=arr(d.mat, d.n_mat, 12, np.float32), point=arr(d.point, d.n_point, 3, np.float32),
                   point_mat=arr(d.point_mat, d.n_point, 1, np.int32))
        if d.has_volume and not d.vol_homogeneous:
            W, H, D = d.vol_res
            out["grid"] = arr(d.grid, W * H * D, 1, np.float32).reshape(D, H, W)
        if d.env_type == 3:
            out["env_map"] = arr(d.env_map, d.env_w * d.env_h, 3, np.float32).reshape(d.env_h, d.env_w, 3)
        return out

    def close(self):
        if getattr(self, "h", None):
            self._lib.ne_scene_file_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SceneReader:
    """SceneReader (src/io/SceneReader.h:8-40) over the library's reader."""

    def __init__(self, filePath=None, absolutePath: bool = False, resources_dir=None, lenient: bool = False):
        self.resources_dir = resources_dir
        self.lenient = lenient
        self._file = None
        if filePath is not None:
            self.loadScene(filePath, absolutePath)

    def loadScene(self, filePath, absolutePath: bool = False):
        if absolutePath or self.resources_dir is None:
            path = str(filePath)
        else:
            path = os.path.join(str(self.resources_dir), str(filePath))
        base = self.resources_dir if self.resources_dir is not None else os.path.dirname(path)
        self._file = SceneFile(path, resources_dir=base, lenient=self.lenient)

    def getScene(self) -> SceneFile:
        return self._file

    def getMainCamera(self) -> Camera:
        return self._file.camera

    def getSettings(self) -> SceneSettings:
        return self._file.settings


def read_vol(path) -> np.ndarray:
    """A .vol grid as a (D, H, W) float32 array (x fastest in memory)."""
    L = lib()
    res = (C.c_int32 * 3)()
    p = os.fsencode(str(path))
    check(L.ne_vol_read(p, res, None, 0))
    W, H, D = res
    g = np.empty(W * H * D, np.float32)
    check(L.ne_vol_read(p, res, g.ctypes.data_as(C.POINTER(C.c_float)), g.size))
    return g.reshape(D, H, W)


def write_vol(path, grid: np.ndarray, comment: str = "") -> None:
    """Writes the text .vol format: "W H D " on line 1, a free second line
    (skipped by readers), then the values, x fastest, each followed by a space
    so that the reference's space-delimited reader keeps all of them."""
    g = np.ascontiguousarray(grid, np.float32)
    D, H, W = g.shape
    with open(path, "w") as f:
        f.write(f"{W} {H} {D} \n{comment}\n")
        flat = g.reshape(-1)
        for k in range(0, flat.size, 4096):
            f.write("".join(f"{float(v)!r} " for v in flat[k:k + 4096]))
        f.write("\n")


def write_pfm(path, rgb: np.ndarray, little_endian: bool = True) -> None:
    """Portable FloatMap of an (H, W, 3) image whose row 0 is the top (PFM
    stores rows bottom to top)."""
    img = np.ascontiguousarray(rgb, np.float32)
    H, W, _ = img.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{W} {H}\n{-1.0 if little_endian else 1.0}\n".encode())
        f.write(img[::-1].astype("<f4" if little_endian else ">f4").tobytes())


def _rgbe(img):
    m = img.max(axis=2)
    e = np.zeros(m.shape, np.int32)
    mant, ex = np.frexp(m)
    ok = m > 1e-32
    e[ok] = ex[ok]
    scale = np.where(ok, 256.0 / np.ldexp(1.0, e), 0.0)
    q = np.zeros(img.shape[:2] + (4,), np.uint8)
    q[..., :3] = np.clip(np.floor(img * scale[..., None]), 0, 255).astype(np.uint8)
    q[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return q


def write_hdr(path, rgb: np.ndarray, rle: bool = True) -> None:
    """Radiance RGBE image of an (H, W, 3) array (row 0 at the top), with
    new-style run-length scanlines when rle (and 8 <= W < 32768)."""
    img = np.asarray(rgb, np.float64)
    H, W, _ = img.shape
    q = _rgbe(img)
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {H} +X {W}\n".encode())
    for y in range(H):
        if not (rle and 8 <= W < 32768):
            out += q[y].tobytes()
            continue
        out += bytes([2, 2, W >> 8, W & 255])
        for c in range(4):
            ch = q[y, :, c]
            x = 0
            while x < W:
                r = 1
                while x + r < W and r < 127 and ch[x + r] == ch[x]:
                    r += 1
                if r >= 3:
                    out += bytes([128 + r, int(ch[x])])
                    x += r
                    continue
                s0 = x
                while x < W and x - s0 < 128 and not (x + 2 < W and ch[x] == ch[x + 1] == ch[x + 2]):
                    x += 1
                out += bytes([x - s0]) + ch[s0:x].tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))


def decode_rgbe(q: np.ndarray) -> np.ndarray:
    """The radiance an RGBE quadruple stands for: mantissa * 2^(e - 136)."""
    q = np.asarray(q)
    f = np.where(q[..., 3] > 0, np.ldexp(1.0, q[..., 3].astype(np.int32) - 136), 0.0)
    return (q[..., :3] * f[..., None]).astype(np.float32)


def write_f32(path, grid: np.ndarray) -> tuple:
    """Writes a raw little-endian float32 grid (x fastest); returns the
    resolution [W, H, D] the scene file's material must state."""
    g = np.ascontiguousarray(grid, np.dtype("<f4"))
    D, H, W = g.shape
    g.tofile(path)
    return (W, H, D)
