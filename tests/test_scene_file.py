"""Scene files (host-side, no GPU): the library's JSON scene reader against the
reference's own test scenes (resources/scenes/testing/*.json, copied under
tests/golden/resources) and the Python scene builder, the .vol / .f32 grid
readers, the OBJ reader and the error behaviour of SceneReader
(src/io/SceneReader.cpp:10-675)."""
import json
from pathlib import Path

import numpy as np
import pytest

from narvalengine_amd import abi
from narvalengine_amd.abi import NarvalError
from narvalengine_amd.scene import Scene, transform
from narvalengine_amd.scene_file import SceneFile, SceneReader, read_vol, write_f32, write_vol

RES = Path(__file__).resolve().parent / "golden" / "resources"

CAMERA = {"position": [0, 1, -3], "lookAt": [0, 2, 0], "up": [0, 1, 0], "speed": 5, "vfov": 45,
          "aperture": 0.0001, "autoFocus": True, "focus": 1}
RENDERER = {"resolution": [64, 32], "spp": 4, "bounces": 6, "mode": "offline", "HDR": False, "toneMapping": False}


def _doc(materials, primitives, camera=CAMERA, renderer=RENDERER):
    return json.dumps({"version": "0.0.1a", "materials": materials, "primitives": primitives, "camera": camera,
                       "renderer": renderer})


def _rect(name, mat, pos, rot, scale):
    return {"name": name, "type": "rectangle", "materialName": mat,
            "transform": {"position": pos, "rotation": rot, "scale": scale}}


def _code(fn):
    with pytest.raises(NarvalError) as e:
        fn()
    return str(e.value)


def test_reference_test_scene_surface_and_light():
    r = SceneReader("scenes/testing/surfaceAndLight.json", False, resources_dir=RES)
    f = r.getScene()
    a = f.arrays()
    # the same scene through the Python builder (an independent FP64 numpy transform)
    s = Scene()
    floor = s.add_glossy((1, 1, 1), 0.95, 0.0)
    light = s.add_emitter((10, 10, 10))
    s.add_rectangle(light, position=(0, 2.0, 0), rotation=(-90, 0, 0), scale=(1, 1, 1))
    s.add_rectangle(floor, position=(0, 0, 0), rotation=(90, 0, 0), scale=(1, 1, 1))
    b = s.arrays()
    np.testing.assert_allclose(a["rect"], b["rect"], rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(a["mat"][a["rect_mat"]], b["mat"][b["rect_mat"]])
    np.testing.assert_allclose(a["rect"][0, 9:12], [0, -1, 0], atol=1e-7)  # the light faces down
    np.testing.assert_allclose(a["rect"][1, 9:12], [0, 1, 0], atol=1e-7)  # the floor faces up
    cam = r.getMainCamera()
    assert cam.position == (0.0, 1.0, -3.0) and cam.look_at == (0.0, 2.0, 0.0) and cam.up == (0.0, 1.0, 0.0)
    assert cam.vfov == 45.0 and cam.aspect == 2.0
    st = r.getSettings()
    assert st.resolution == (600, 300) and st.spp == 1 and st.bounces == 6 and st.hdr is False
    assert f.realtime and f.warnings == []
    d, _ = f.desc()
    assert d.has_volume == 0 and d.has_sun == 0 and d.env_type == abi.NE_ENV_NONE


def test_reference_test_scene_single_voxel_volume():
    f = SceneReader("scenes/testing/singleVoxelVolume.json", resources_dir=RES).getScene()
    d, _ = f.desc()
    assert d.has_volume == 1 and d.vol_homogeneous == 0 and tuple(d.vol_res) == (1, 1, 1)
    assert f.arrays()["grid"].reshape(-1).tolist() == [1.0]
    assert d.vol_mult == 100.0 and d.vol_g == 0.0
    np.testing.assert_allclose(d.vol_sigma_s, 1.1, rtol=1e-7)
    np.testing.assert_allclose(d.vol_sigma_a, 0.01, rtol=1e-6)
    M = transform((0, 0.5, 0), (0, 0, 0), (1, 1, 1))
    np.testing.assert_allclose(np.array(d.vol_xf).reshape(4, 4), M, atol=1e-7)
    np.testing.assert_allclose(np.array(d.vol_inv).reshape(4, 4), np.linalg.inv(M), atol=1e-7)
    assert d.maj_res == 1  # capped at the grid size
    assert f.arrays()["rect"].shape == (2, 12)


def test_vol_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(1)
    g = rng.uniform(0, 3, (5, 4, 7)).astype(np.float32)
    g[0, 0, 0] = 1e-30
    g[1, 2, 3] = 123456.789
    p = tmp_path / "g.vol"
    write_vol(p, g, comment="test grid")
    np.testing.assert_array_equal(read_vol(p), g)  # bit-exact
    # newline-separated values and no trailing space are accepted too
    q = tmp_path / "h.vol"
    q.write_text("2 1 1\n\n0.25\n0.5")
    np.testing.assert_array_equal(read_vol(q), np.array([[[0.25, 0.5]]], np.float32))
    bad = tmp_path / "bad.vol"
    bad.write_text("2 2 2 \n\n1 2 3 \n")
    assert "3 values" in _code(lambda: read_vol(bad))
    bad.write_text("2 2\n\n1 2 3 4 \n")
    assert "W H D" in _code(lambda: read_vol(bad))
    bad.write_text("1 1 1\n\nabc \n")
    assert "invalid value" in _code(lambda: read_vol(bad))
    assert "Couldn't read" in _code(lambda: read_vol(tmp_path / "missing.vol"))


def _write_obj(path):
    # a unit cube: quads, one polygon of 5 vertices, v/vt/vn and negative indices
    path.write_text("""# cube
o cube
v -0.5 -0.5 -0.5
v 0.5 -0.5 -0.5
v 0.5 0.5 -0.5
v -0.5 0.5 -0.5
v -0.5 -0.5 0.5
v 0.5 -0.5 0.5
v 0.5 0.5 0.5
v -0.5 0.5 0.5
vt 0 0
vn 0 0 -1
usemtl x
f 1/1/1 2/1/1 3/1/1 4/1/1
f 5//1 8//1 7//1 6//1
f -8 -4 -1 -5
f 2 6 7 3
f 4 3 7 8
f 1 5 6 2 1
""")
    v = np.array([[-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5], [-0.5, 0.5, -0.5],
                  [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]])
    faces = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 7, 3], [1, 5, 6, 2], [3, 2, 6, 7], [0, 4, 5, 1, 0]]
    tris = [[f[0], f[i], f[i + 1]] for f in faces for i in range(1, len(f) - 1)]
    return v, np.array(tris)


def test_obj_mesh_with_instance_transform(tmp_path):
    (tmp_path / "models").mkdir()
    v, t = _write_obj(tmp_path / "models" / "cube.obj")
    pos, rot, scale = [0.2, 0.3, -0.1], [30, -20, 45], [1.5, 0.5, 2.0]
    doc = _doc([{"name": "m", "type": "microfacet", "roughness": 0.4, "metallic": 0.1, "albedo": [0.2, 0.3, 0.4]}],
               [{"name": "cube", "type": "obj", "path": "models/cube.obj", "materialName": "m",
                 "transform": {"position": pos, "rotation": rot, "scale": scale}}])
    f = SceneFile(text=doc, resources_dir=tmp_path)
    a = f.arrays()
    M = transform(pos, rot, scale)
    w = v @ M[:3, :3].T + M[:3, 3]
    np.testing.assert_allclose(a["tri_v"], w[t].reshape(-1, 9), rtol=1e-6, atol=1e-6)
    assert (a["tri_mat"] == 0).all() and len(a["tri_mat"]) == len(t) == 13
    np.testing.assert_allclose(a["mat"][0, :6], [0, 0.2, 0.3, 0.4, 0.4, 0.1], rtol=1e-7)
    # a face index past the vertex list is a parse error with its line
    (tmp_path / "models" / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    doc_bad = doc.replace("models/cube.obj", "models/bad.obj")
    msg = _code(lambda: SceneFile(text=doc_bad, resources_dir=tmp_path))
    assert "line 4" in msg and "out of range" in msg


def test_lights_spheres_volumes_and_replacement(tmp_path):
    grid = np.random.default_rng(2).uniform(0, 1, (6, 5, 4)).astype(np.float32)
    W, H, D = write_f32(tmp_path / "g.f32", grid)
    mats = [
        {"name": "e", "type": "emitter", "albedo": [1, 2, 3]},
        {"name": "e", "type": "emitter", "albedo": [4, 5, 6]},  # replaceMaterial: the last one wins
        {"name": "sun", "type": "directionalLight", "albedo": [2, 2, 1.5], "position": [1, 4, -2]},
        {"name": "vol", "type": "volume", "scattering": [0.9, 0.8, 0.7], "absorption": [0.1, 0.1, 0.1],
         "phaseFunction": "henyey-greenstein", "g": 0.6, "density": 7, "path": "g.f32", "resolution": [W, H, D]},
        {"name": "gray", "type": "microfacet", "roughness": 0.5, "metallic": 0, "albedo": [0.5, 0.5, 0.5],
         "normalMap": [0.5, 0.5, 1.0]},
    ]
    prims = [
        {"name": "ball", "type": "sphere", "radius": 0.25, "materialName": "e",
         "transform": {"position": [0.1, 0.2, 0.3]}},
        {"name": "ghost", "type": "sphere", "radius": 0.5, "materialName": "gray", "collision": False,
         "transform": {"position": [0, 0, 0]}},
        {"name": "sunpoint", "type": "point", "materialName": "sun",
         "transform": {"position": [0, 0, 0], "rotation": [0, 0, 0], "scale": [1, 1, 1]}},
        {"name": "cloud", "type": "volume", "materialName": "vol",
         "transform": {"position": [0, 1, 0], "rotation": [0, 30, 0], "scale": [2, 1, 1]}},
    ]
    f = SceneFile(text=_doc(mats, prims), resources_dir=tmp_path, maj_res=3)
    a = f.arrays()
    d, _ = f.desc()
    np.testing.assert_array_equal(a["sph"], np.float32([[0.1, 0.2, 0.3, 0.25]]))  # the grey ghost is dropped
    np.testing.assert_array_equal(a["mat"][a["sph_mat"][0], 6:9], [4, 5, 6])
    assert d.has_sun == 1
    np.testing.assert_allclose(d.sun_dir, -np.array([1, 4, -2]) / np.sqrt(21), rtol=1e-6)
    np.testing.assert_allclose(d.sun_E, [2, 2, 1.5])
    assert tuple(d.vol_res) == (W, H, D) == (4, 5, 6) and d.maj_res == 3
    np.testing.assert_array_equal(a["grid"], grid)
    assert d.vol_mult == 7 and abs(d.vol_g - 0.6) < 1e-7
    # a volume material without a path is a homogeneous medium
    mats[3] = {"name": "vol", "type": "volume", "scattering": [1, 1, 1], "absorption": [0, 0, 0],
               "phaseFunction": "isotropic", "density": 2.5}
    d2, _ = SceneFile(text=_doc(mats, prims)).desc()
    assert d2.has_volume == 1 and d2.vol_homogeneous == 1 and d2.vol_mult == 2.5 and d2.vol_g == 0.0


def test_errors_follow_the_reference_reader(tmp_path):
    ok_mat = {"name": "m", "type": "microfacet", "roughness": 0.5, "metallic": 0, "albedo": [1, 1, 1]}
    rect = _rect("r", "m", [0, 0, 0], [0, 0, 0], [1, 1, 1])
    parse = lambda text, **kw: SceneFile(text=text, **kw)
    assert "version" in _code(lambda: parse(json.dumps({"materials": [], "primitives": []})))
    assert "Malformatted" in _code(lambda: parse("[1, 2]"))
    assert "line 2" in _code(lambda: parse('{"version": "0.0.1a",\n "materials": [1,],}'))
    assert "Material is missing the property name." in _code(lambda: parse(_doc([{"type": "emitter"}], [])))
    assert "Material m is missing the property type." in _code(lambda: parse(_doc([{"name": "m"}], [])))
    assert "Invalid material type" in _code(lambda: parse(_doc([{"name": "d", "type": "diffuse"}], [])))
    assert "missing the property roughness" in _code(
        lambda: parse(_doc([{"name": "m", "type": "microfacet", "metallic": 0, "albedo": [1, 1, 1]}], [])))
    box = {"name": "b", "type": "box", "transform": {"position": [0, 0, 0]}}
    assert "Invalid primitive type" in _code(lambda: parse(_doc([ok_mat], [box])))
    assert "unknown material" in _code(lambda: parse(_doc([ok_mat], [_rect("r", "nope", [0, 0, 0], [0, 0, 0],
                                                                           [1, 1, 1])])))
    assert "missing the property camera" in _code(
        lambda: parse(json.dumps({"version": "0.0.1a", "materials": [ok_mat], "primitives": [rect]})))
    assert "Couldn't read" in _code(lambda: SceneFile(tmp_path / "none.json"))
    # valid input the backend cannot represent: an error, or a listed substitution when lenient
    tex = dict(ok_mat, albedo="textures/wood.png")
    assert "not supported" in _code(lambda: parse(_doc([tex], [rect])))
    f = SceneFile(text=_doc([tex], [rect]), lenient=True)
    assert len(f.warnings) == 1 and "grey" in f.warnings[0]
    np.testing.assert_array_equal(f.arrays()["mat"][0, 1:4], [0.5, 0.5, 0.5])
    vdb = {"name": "v", "type": "volume", "scattering": [1, 1, 1], "absorption": [0, 0, 0], "phaseFunction": "hg",
           "g": 0.5, "density": 1, "path": "vdb/cloud.vdb"}
    assert "OpenVDB" in _code(lambda: parse(_doc([vdb], [])))
    vp = {"name": "vp", "type": "volume", "materialName": "v",
          "transform": {"position": [0, 0, 0], "rotation": [0, 0, 0], "scale": [1, 1, 1]}}
    assert "could not be read" in _code(lambda: parse(_doc([vdb], [vp]), lenient=True))
    homog = dict(vdb, name="h")
    homog.pop("path")
    vp2 = dict(vp, materialName="h")
    assert "one volume per scene" in _code(lambda: parse(_doc([homog], [vp2, dict(vp2, name="other")])))
    gltf = {"name": "g", "type": "gltf", "path": "x.gltf", "materialName": "m",
            "transform": {"position": [0, 0, 0], "rotation": [0, 0, 0], "scale": [1, 1, 1]}}
    assert "glTF" in _code(lambda: parse(_doc([ok_mat], [gltf])))
    pl = {"name": "p", "type": "point", "materialName": "e",
          "transform": {"position": [0.5, 2, -1], "rotation": [0, 0, 0], "scale": [1, 1, 1]}}
    f = parse(_doc([{"name": "e", "type": "emitter", "albedo": [3, 2, 1]}], [pl]))
    a = f.arrays()
    np.testing.assert_array_equal(a["point"], np.float32([[0.5, 2, -1]]))
    np.testing.assert_array_equal(a["mat"][a["point_mat"][0], 6:9], [3, 2, 1])


def test_scene_file_drives_the_oracle_like_the_builder(tmp_path):
    """A scene written as a file and the same scene built in Python give the
    oracle identical inputs, hence identical images."""
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    import oracle as O
    mats = [{"name": "floor", "type": "microfacet", "roughness": 0.6, "metallic": 0.0, "albedo": [0.7, 0.6, 0.5]},
            {"name": "lamp", "type": "emitter", "albedo": [5, 5, 5]}]
    prims = [_rect("floor", "floor", [0, 0, 0], [90, 0, 0], [4, 4, 1]),
             _rect("lamp", "lamp", [0.2, 1.5, 0.1], [-90, 10, 0], [0.5, 0.5, 1]),
             {"name": "ball", "type": "sphere", "radius": 0.3, "materialName": "floor",
              "transform": {"position": [0, 0.3, 0]}}]
    cam = dict(CAMERA, position=[0, 1, -2.5], lookAt=[0, 0.3, 0])
    f = SceneFile(text=_doc(mats, prims, camera=cam))
    s = Scene()
    fl = s.add_glossy((0.7, 0.6, 0.5), 0.6, 0.0)
    lp = s.add_emitter((5, 5, 5))
    s.add_rectangle(fl, position=(0, 0, 0), rotation=(90, 0, 0), scale=(4, 4, 1))
    s.add_rectangle(lp, position=(0.2, 1.5, 0.1), rotation=(-90, 10, 0), scale=(0.5, 0.5, 1))
    s.add_sphere((0, 0.3, 0), 0.3, fl)
    a, b = f.arrays(), s.arrays()
    for k in ("rect", "sph", "mat"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-6, atol=1e-7)
    ia = O.World(f).render(f.camera, 16, 8, 4, 4, seed=3)
    ib = O.World(s).render(f.camera, 16, 8, 4, 4, seed=3)
    np.testing.assert_allclose(ia, ib, rtol=1e-4, atol=1e-5)
    assert ia.mean() > 0


def test_environment_images(tmp_path):
    """infiniteAreaLight on a point primitive: Radiance .hdr (run-length and
    flat scanlines) and .pfm (either byte order) maps, row 0 at the top."""
    from narvalengine_amd.scene_file import _rgbe, decode_rgbe, write_hdr, write_pfm
    rng = np.random.default_rng(7)
    img = rng.gamma(0.7, 2.0, (6, 40, 3)).astype(np.float32)
    img[1, 5:30] = (3.0, 2.0, 1.0)  # long runs for the run-length coder
    img[4] = 0.0
    expect_hdr = decode_rgbe(_rgbe(img.astype(np.float64)))
    # RGBE keeps 8 mantissa bits relative to the brightest channel of a pixel
    assert np.all(np.abs(expect_hdr - img) <= img.max(axis=2, keepdims=True) / 128 + 1e-30)
    point = {"name": "sky", "type": "point", "materialName": "env",
             "transform": {"position": [0, 0, 0], "rotation": [0, 0, 0], "scale": [1, 1, 1]}}
    cases = []
    for name, writer, expect in (("rle.hdr", lambda p: write_hdr(p, img, rle=True), expect_hdr),
                                 ("flat.hdr", lambda p: write_hdr(p, img, rle=False), expect_hdr),
                                 ("le.pfm", lambda p: write_pfm(p, img, True), img),
                                 ("be.pfm", lambda p: write_pfm(p, img, False), img)):
        writer(tmp_path / name)
        mats = [{"name": "env", "type": "infiniteAreaLight", "path": name}]
        f = SceneFile(text=_doc(mats, [point]), resources_dir=tmp_path)
        d, _ = f.desc()
        assert d.env_type == abi.NE_ENV_MAP and (d.env_w, d.env_h) == (40, 6)
        np.testing.assert_array_equal(f.arrays()["env_map"], expect)
        cases.append(name)
    assert len(cases) == 4
    # a material that no point uses is loaded but lights nothing
    d, _ = SceneFile(text=_doc([{"name": "env", "type": "infiniteAreaLight", "path": "le.pfm"}], []),
                     resources_dir=tmp_path).desc()
    assert d.env_type == abi.NE_ENV_NONE
    mats = [{"name": "env", "type": "infiniteAreaLight", "path": "sky.png"}]
    assert ".hdr or .pfm" in _code(lambda: SceneFile(text=_doc(mats, [point])))
    (tmp_path / "bad.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 2 +X 2\n" + bytes(16))
    mats = [{"name": "env", "type": "infiniteAreaLight", "path": "bad.hdr"}]
    assert "orientation" in _code(lambda: SceneFile(text=_doc(mats, [point]), resources_dir=tmp_path))
