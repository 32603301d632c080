"""Output writers (ne_image_write), host-only: the reference's output.ppm
layout (OfflineEngine::coreLoop, src/core/OfflineEngine.cpp:80-137), PNG read
back by Pillow, PFM read back by the library's own PFM reader (the
environment loader), and EXR parsed by the OpenEXR file layout."""
import json
import struct

import numpy as np
import pytest
from PIL import Image

from narvalengine_amd.abi import NE_IMAGE_EXR, NE_IMAGE_PFM, NE_IMAGE_PNG, NE_IMAGE_PPM16, NarvalError
from narvalengine_amd.engine import write_image
from narvalengine_amd.scene_file import SceneFile


def _img(H=7, W=11, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.2, 1.2, (H, W, 3)).astype(np.float32)
    a[0, 0] = (0.0, 1.0, 0.5)
    a[1, 1] = (np.nan, 2.0, -1.0)
    return a


def test_ppm16_is_the_reference_layout(tmp_path):
    img = _img()
    H, W, _ = img.shape
    p = tmp_path / "out.ppm"
    write_image(p, img, NE_IMAGE_PPM16)
    b = p.read_bytes()
    head = f"P6\n{W} {H}\n65535\n".encode()
    assert b[:len(head)] == head and len(b) == len(head) + W * H * 6
    v = np.frombuffer(b[len(head):], ">u2").reshape(-1, 3)
    # pixels from the last index to the first; value * 65535 truncated after clamping
    flat = np.nan_to_num(img.reshape(-1, 3), nan=0.0)
    expect = (np.clip(flat, 0, 1) * np.float32(65535)).astype(np.uint16)[::-1]
    np.testing.assert_array_equal(v, expect)
    assert tuple(v[-1]) == (0, 65535, 32767)  # the first pixel comes last


def test_png_decodes_with_pillow(tmp_path):
    img = _img(H=300, W=230)  # > 65535 raw bytes: several stored deflate blocks
    p = tmp_path / "out.png"
    write_image(p, img, NE_IMAGE_PNG)
    got = np.asarray(Image.open(p).convert("RGB"))
    expect = (np.clip(np.nan_to_num(img, nan=0.0), 0, 1) * np.float32(255)).astype(np.uint8)
    np.testing.assert_array_equal(got, expect)


def test_pfm_round_trips_through_the_environment_loader(tmp_path):
    img = np.abs(_img(seed=3))
    img[1, 1] = (0.5, 2.0, 1.0)
    write_image(tmp_path / "sky.pfm", img, NE_IMAGE_PFM)
    doc = {"version": "0.0.1a", "materials": [{"name": "env", "type": "infiniteAreaLight", "path": "sky.pfm"}],
           "primitives": [{"name": "p", "type": "point", "materialName": "env",
                           "transform": {"position": [0, 0, 0], "rotation": [0, 0, 0], "scale": [1, 1, 1]}}],
           "camera": {"position": [0, 0, -1], "lookAt": [0, 0, 0], "up": [0, 1, 0], "speed": 1, "vfov": 45,
                      "aperture": 0, "focus": 1},
           "renderer": {"resolution": [4, 4], "spp": 1, "bounces": 1, "mode": "offline", "HDR": True}}
    f = SceneFile(text=json.dumps(doc), resources_dir=tmp_path)
    np.testing.assert_array_equal(f.arrays()["env_map"], img)


def _read_exr(path):
    """Single-part scanline OpenEXR without compression, FLOAT channels."""
    b = path.read_bytes()
    magic, version = struct.unpack_from("<ii", b, 0)
    assert magic == 20000630 and version & 0xFF == 2
    p, attrs = 8, {}
    while b[p] != 0:
        e = b.index(b"\0", p)
        name = b[p:e].decode()
        e2 = b.index(b"\0", e + 1)
        typ = b[e + 1:e2].decode()
        (size,) = struct.unpack_from("<i", b, e2 + 1)
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + size])
        p = e2 + 5 + size
    p += 1
    chans, q, raw = [], 0, attrs["channels"][1]
    while raw[q] != 0:
        e = raw.index(b"\0", q)
        ptype = struct.unpack_from("<i", raw, e + 1)[0]
        chans.append((raw[q:e].decode(), ptype))
        q = e + 1 + 16
    assert attrs["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{H}Q", b, p)
    out = np.zeros((H, W, len(chans)), np.float32)
    for off in offsets:
        y, size = struct.unpack_from("<ii", b, off)
        assert size == W * len(chans) * 4
        data = np.frombuffer(b, "<f4", W * len(chans), off + 8).reshape(len(chans), W)
        out[y - y0] = data.T
    return [c[0] for c in chans], [c[1] for c in chans], out


def test_exr_layout(tmp_path):
    img = _img(seed=5) * 40.0
    p = tmp_path / "out.exr"
    write_image(p, img, NE_IMAGE_EXR)
    names, types, data = _read_exr(p)
    assert names == ["B", "G", "R"] and types == [2, 2, 2]  # FLOAT channels in alphabetical order
    np.testing.assert_array_equal(data[..., ::-1], img)


def test_writer_errors(tmp_path):
    img = _img()
    with pytest.raises(NarvalError, match="unknown image format"):
        write_image(tmp_path / "x", img, 9)
    with pytest.raises(NarvalError, match="cannot open"):
        write_image(tmp_path / "no" / "such" / "dir.png", img, NE_IMAGE_PNG)
