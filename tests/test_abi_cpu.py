"""CPU checks of the C ABI library: it loads, exports every symbol declared in
include/ne_gpu.h, its structs match the Python mirror, the host-side BVH is
well formed and the synthetic-input generators are deterministic."""
import ctypes as C
import hashlib
import json
import re
from pathlib import Path

import numpy as np
import pytest

import scenes
from narvalengine_amd import abi, configs

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"


def declared_symbols():
    txt = (ROOT / "include" / "ne_gpu.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ne_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    bound = {s[0] for s in abi.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_struct_layouts():
    assert abi.lib().ne_sizeof_scene_desc() == C.sizeof(abi.SceneDesc)
    import oracle as O
    assert O.lib().neo_sizeof_scene_desc() == C.sizeof(abi.SceneDesc)
    assert O.lib().neo_sizeof_camera_desc() == C.sizeof(abi.CameraDesc)


def test_context_without_device_fails_cleanly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = C.c_void_p()
    st = abi.lib().ne_ctx_create(0, C.byref(h))
    assert st == 4  # NE_ERR_NO_DEVICE
    assert not h.value
    assert b"device" in abi.lib().ne_last_error(None)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_generators_are_deterministic():
    ref = json.loads((GOLDEN / "generator_checksums.json").read_text())
    g = np.zeros((32, 32, 32), np.float32)
    assert abi.lib().ne_gen_fbm_grid(32, 0, 4.0, g.ctypes.data_as(C.POINTER(C.c_float))) == 0
    assert _sha(g) == ref["fbm32_seed0"]
    c = np.zeros((32, 32, 32), np.float32)
    assert abi.lib().ne_gen_cumulus_grid(32, 2, c.ctypes.data_as(C.POINTER(C.c_float))) == 0
    assert _sha(c) == ref["cumulus32_seed2"]
    v, t = configs.icosphere(2, 1, 0.1)
    assert len(t) == 20 * 4 ** 2 and len(v) == 10 * 4 ** 2 + 2
    assert _sha(v) == ref["ico2_v"] and _sha(t) == ref["ico2_t"]
    # closed-form properties: density in [0, 1], icosphere edges shared by two faces
    assert g.min() >= 0 and g.max() <= 1 and 0.2 < (g > 0).mean() < 0.8
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()


def _bvh(scene):
    d, keep = scene.desc()
    nn, nr, dep = C.c_uint32(), C.c_uint32(), C.c_uint32()
    L = abi.lib()
    assert L.ne_bvh_build(C.byref(d), None, None, C.byref(nn), C.byref(nr), C.byref(dep)) == 0
    nodes = np.zeros((nn.value, 16), np.float32)
    refs = np.zeros(nr.value, np.uint32)
    assert L.ne_bvh_build(C.byref(d), nodes.ctypes.data_as(C.POINTER(C.c_float)),
                          refs.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(nn), C.byref(nr), C.byref(dep)) == 0
    return nodes, refs, dep.value, keep


@pytest.mark.parametrize("which", ["soup", "box"])
def test_bvh_is_well_formed(which):
    s = scenes.triangle_soup() if which == "soup" else scenes.small_surface(subdiv=4)
    nodes, refs, depth, a = _bvh(s)
    words = nodes.view(np.uint32)
    ints = nodes[:, 12:].view(np.int32)  # child[4]
    n_prims = len(a["tri_v"]) + len(a["sph"]) + len(a["rect"])
    assert len(refs) == n_prims and depth <= 60
    # every primitive appears in exactly one leaf slot
    key = (refs >> 30).astype(np.int64) * (1 << 30) + (refs & 0x3FFFFFFF)
    assert len(np.unique(key)) == n_prims

    def prim_box(r):
        ty, ix = int(r >> 30), int(r & 0x3FFFFFFF)
        if ty == 0:
            v = a["tri_v"][ix].reshape(3, 3)
            return v.min(0), v.max(0)
        if ty == 1:
            c, rad = a["sph"][ix, :3], a["sph"][ix, 3]
            return c - rad, c + rad
        p = a["rect"][ix]
        cs = np.array([p[:3], p[:3] + p[3:6], p[:3] + p[6:9], p[:3] + p[3:6] + p[6:9]])
        return cs.min(0), cs.max(0)

    def child_box(k, c):
        # FP32 decode as the traversal does it: origin + byte * 2^e per plane
        origin = nodes[k, 0:3]
        scale = np.array([np.ldexp(np.float32(1), int((words[k, 3] >> (8 * a)) & 255) - 127) for a in range(3)],
                         np.float32)
        q = lambda w: np.float32((int(words[k, w]) >> (8 * c)) & 255)
        lo = np.array([origin[a] + q(4 + a) * scale[a] for a in range(3)], np.float32)
        hi = np.array([origin[a] + q((7, 8, 9)[a]) * scale[a] for a in range(3)], np.float32)
        return lo, hi

    visited = np.zeros(len(nodes), bool)
    slots = 0

    def subtree(k):
        """Visits node k; returns the union of its primitives' true boxes after
        checking that every decoded child box contains its subtree's primitives."""
        nonlocal slots
        assert not visited[k]
        visited[k] = True
        assert (ints[k, :4] >= 0).sum() >= (1 if k == 0 else 2)  # a wide node has 2..4 children
        ulo, uhi = np.full(3, np.inf), np.full(3, -np.inf)
        for side in range(4):
            # leaf meta byte: 0x80 | primitive type << 4 | count; leaves hold one type
            idx, meta = ints[k, side], (int(words[k, 10]) >> (8 * side)) & 255
            cnt, leaf_type = (meta & 15) if meta & 0x80 else 0, (meta >> 4) & 3
            if idx < 0:
                continue
            lo, hi = child_box(k, side)
            if cnt > 0:
                assert 0 <= idx and idx + cnt <= len(refs)
                slots += cnt
                boxes = []
                for r in refs[idx:idx + cnt]:
                    assert int(r >> 30) == leaf_type
                    boxes.append(prim_box(r))
                clo, chi = np.min([b[0] for b in boxes], 0), np.max([b[1] for b in boxes], 0)
            else:
                assert 0 < idx < len(nodes)
                clo, chi = subtree(int(idx))
            assert (clo >= lo).all() and (chi <= hi).all()  # conservative quantization
            ulo, uhi = np.minimum(ulo, clo), np.maximum(uhi, chi)
        return ulo, uhi

    subtree(0)
    assert visited.all() and slots == len(refs)


def test_invalid_scene_is_rejected_without_device():
    from narvalengine_amd.scene import Scene
    s = Scene()
    m = s.add_glossy((0.5, 0.5, 0.5), 0.5)
    s.add_sphere((0, 0, 0), 0.0, m)
    d, keep = s.desc()
    nn = C.c_uint32()
    assert abi.lib().ne_bvh_build(C.byref(d), None, None, C.byref(nn), None, None) == abi.lib().ne_bvh_build(
        C.byref(d), None, None, None, None, None) == 1
    s2 = scenes.small_volume(res=8, maj_res=2)
    g = np.array(s2.volume["grid"])
    g[1, 2, 3] = np.nan
    s2.volume["grid"] = g
    d2, keep2 = s2.desc()
    assert abi.lib().ne_bvh_build(C.byref(d2), None, None, None, None, None) == 1
    assert b"finite" in abi.lib().ne_last_error(None)


@pytest.mark.parametrize("bad", ["vertex", "radius", "centre", "corner"])
def test_non_finite_geometry_is_rejected_before_the_bvh_build(bad):
    """A primitive with an infinite or NaN coordinate has no conservative
    box, so the scene is rejected instead of being built with boxes that miss it."""
    from narvalengine_amd.scene import Scene
    s = Scene()
    m = s.add_glossy((0.5, 0.5, 0.5), 0.5)
    s.add_sphere((0, 0, 0), 0.5, m)
    s.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.array([[0, 1, 2]]), m)
    s.add_rectangle(m, position=(0, 0, 1), rotation=(0, 0, 0), scale=(1, 1, 1))
    if bad == "vertex":
        s.tri_v[-1][0, 4] = np.inf
    elif bad == "radius":
        s.sph[-1][3] = np.inf
    elif bad == "centre":
        s.sph[-1][1] = np.nan
    d, keep = s.desc()
    if bad == "corner":
        keep_rect = np.array(np.ctypeslib.as_array(d.rect, shape=(12,)))
        keep_rect[5] = -np.inf
        d.rect = keep_rect.ctypes.data_as(C.POINTER(C.c_float))
    assert abi.lib().ne_bvh_build(C.byref(d), None, None, None, None, None) == 1
    assert b"finite" in abi.lib().ne_last_error(None)


def test_path_id_divmod_method_is_exact():
    """The wavefront's path id -> (sample, pixel) split takes floor of the
    correctly rounded double quotient and one correction step each way (the
    quotient is within 2^-20 of the true one for quotients below 2^32).  The
    same IEEE operations in numpy against Python's exact divmod, over the
    stated range: divisors up to 2^30 (rect pixel counts) and quotients below
    2^32, at and around exact multiples."""
    rng = np.random.default_rng(5)

    def method(i, n):
        q = int(np.floor(np.float64(i) / np.float64(n)))
        r = i - q * n
        if r < 0:
            q, r = q - 1, r + n
        if r >= n:
            q, r = q + 1, r - n
        return q, r

    divisors = [1, 2, 3, 7, 64, 255, 1920 * 1080, 32768 * 32767, 2 ** 30 - 1, 2 ** 30]
    divisors += [int(x) for x in rng.integers(1, 2 ** 30, 20)]
    for n in divisors:
        ids = [int(x) for x in rng.integers(0, n * 2 ** 32, 200, dtype=np.uint64)]
        for k in (1, 2, 3, 2 ** 20 + 1, 2 ** 31, 2 ** 32 - 1):
            ids += [k * n - 1, k * n, k * n + 1]
        for i in ids:
            if 0 <= i < n * 2 ** 32:
                assert method(i, n) == divmod(i, n), (i, n)
