"""GPU parity: device code (through the C ABI) against the oracle on the same
seeded inputs.  Bit-exact for hit/miss, primitive ids and RNG words; t within
1e-5 relative; pdf/eval within 1e-5 relative; images statistically equal."""
import numpy as np
import pytest

import oracle as O
import scenes
from conftest import print_case
from narvalengine_amd import abi
from narvalengine_amd.abi import NE_FP32, NE_FP64
from narvalengine_amd.scene import Scene

pytestmark = pytest.mark.gpu
RTOL = 1e-5


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


# ---- RNG ---------------------------------------------------------------------
@pytest.mark.parametrize("seed,pixel,sample,job", [(0, 0, 0, 0), (0x0123456789abcdef, 1234567, 99, 7),
                                                  (2 ** 64 - 1, 2 ** 32 - 1, 2 ** 32 - 1, 5)])
def test_rng_stream_bit_exact(ctx, seed, pixel, sample, job):
    print_case(f"rng {seed} {pixel} {sample} {job}")
    g = ctx.test_rng(seed, pixel, sample, job, 37)
    o = O.rng_uniforms(seed, pixel, sample, job, 37)
    np.testing.assert_array_equal(g, o)


# ---- primitives (reference semantics) --------------------------------------
def _prim_cases(kind, n, seed):
    rng = np.random.default_rng(seed)
    rays = scenes.random_rays(n, seed=seed, origin_box=1.5)
    if kind == abi.NE_TEST_AABB:
        lo = rng.uniform(-1, 0.5, (n, 3))
        prims = np.concatenate([lo, lo + rng.uniform(0.01, 1.0, (n, 3))], axis=1)
        # aim half of the rays at the box centres so both outcomes are frequent
        c = (prims[:, :3] + prims[:, 3:]) / 2
    elif kind == abi.NE_TEST_TRIANGLE:
        c = rng.uniform(-0.5, 0.5, (n, 3))
        prims = (c[:, None, :] + rng.normal(0, 0.3, (n, 3, 3))).reshape(n, 9)
        c = prims.reshape(n, 3, 3).mean(axis=1)
    elif kind == abi.NE_TEST_SPHERE:
        c = rng.uniform(-0.5, 0.5, (n, 3))
        prims = np.concatenate([c, rng.uniform(0.05, 0.6, (n, 1))], axis=1)
    else:
        # reference rectangle: vertexData[0], vertexData[1] corners of an
        # axis-aligned quad, normal along the flat axis
        c = rng.uniform(-0.5, 0.5, (n, 3))
        axis = rng.integers(0, 3, n)
        a = c - rng.uniform(0.05, 0.5, (n, 3))
        b = c + rng.uniform(0.05, 0.5, (n, 3))
        nrm = np.zeros((n, 3))
        for i in range(n):
            b[i, axis[i]] = a[i, axis[i]] = c[i, axis[i]]
            nrm[i, axis[i]] = -1.0
        prims = np.concatenate([a, b, nrm], axis=1)
    aim = rng.random(n) < 0.5
    d = c - rays[:, :3]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[aim, 3:] = d[aim]
    return rays.astype(np.float32), prims.astype(np.float32)


@pytest.mark.parametrize("kind", [abi.NE_TEST_AABB, abi.NE_TEST_TRIANGLE, abi.NE_TEST_SPHERE, abi.NE_TEST_RECTANGLE])
def test_primitive_parity(ctx, kind):
    print_case(f"primitive kind {kind}")
    n = 20000
    rays, prims = _prim_cases(kind, n, seed=100 + kind)
    hit, t0, t1 = ctx.test_primitive(kind, rays, prims)
    if kind == abi.NE_TEST_AABB:
        oh, ot0, ot1, _ = O.aabb(rays, prims)
    elif kind == abi.NE_TEST_TRIANGLE:
        oh, ot0, _ = O.triangle(rays, prims)
        ot1 = ot0
    elif kind == abi.NE_TEST_SPHERE:
        oh, ot0, ot1 = O.sphere(rays, prims)
    else:
        oh, ot0 = O.rectangle(rays, prims)
        ot1 = ot0
    np.testing.assert_array_equal(hit, oh)
    assert 0.1 < hit.mean() < 0.9
    m = oh == 1
    assert (_rel(t0[m], ot0[m]) <= RTOL).all()
    assert (_rel(t1[m], ot1[m]) <= RTOL).all()


def test_reference_goldens_on_device(ctx):
    print_case("goldens on device")
    tri = np.array([0, 0, 0, 0, 1, 0, 1, 0, 0], np.float32)
    rays = np.array([[0, 0, 0, 0, 0, 1], [0, 0, -1, 0, 0, 1], [0, -1, 0, 0, 1, 0], [0.25, 0.25, -1, 0, 0, 1],
                     [0, 0, 1, 0, 0, 1]], np.float32)
    hit, t, _ = ctx.test_primitive(abi.NE_TEST_TRIANGLE, rays, np.tile(tri, (5, 1)))
    assert list(hit) == [1, 1, 0, 1, 0]
    assert t[0] == 0.0 and t[1] == 1.0 and t[3] == 1.0
    box = np.array([-0.5, -0.5, -0.5, 0.5, 0.5, 0.5], np.float32)
    rays = np.array([[0, 0, -1, 0, 0, 1], [0, 0, -0.5, 0, 0, 1], [0, -0.5, -0.5, 0, 0, 1], [0, 0, 0, 0, 0, 1],
                     [0, 0.55, 0, 0, 0, 1]], np.float32)
    hit, tn, tf = ctx.test_primitive(abi.NE_TEST_AABB, rays, np.tile(box, (5, 1)))
    assert list(hit) == [1, 1, 1, 1, 0]
    np.testing.assert_allclose(tn[:4], [0.5, 0, 0, -0.5], atol=1e-3)
    np.testing.assert_allclose(tf[:4], [1.5, 1, 1, 0.5], atol=1e-3)
    assert tn[4] == np.inf and tf[4] == -np.inf


# ---- phase function / BSDF ---------------------------------------------------
def test_hg_and_ggx_parity(ctx):
    print_case("hg/ggx")
    rng = np.random.default_rng(7)
    n = 20000
    a = rng.normal(size=(n, 3)).astype(np.float32)
    b = rng.normal(size=(n, 3)).astype(np.float32)
    g = rng.uniform(-0.95, 0.95, n).astype(np.float32)
    got = ctx.test_bsdf(abi.NE_TEST_HG_EVAL, n, [a, b, g], 1)[:, 0]
    exp = np.zeros(n, np.float32)
    # oracle HG eval takes a scalar g: evaluate per element
    for i in range(n):
        out = np.zeros(1, np.float32)
        O.lib().neo_hg_eval(1, float(g[i]), O.fp(O.f(a[i:i + 1])), O.fp(O.f(b[i:i + 1])), O.fp(out))
        exp[i] = out[0]
    assert (_rel(got, exp) <= RTOL).all()
    # HG sampler (reference local frame)
    u = rng.random((n, 2)).astype(np.float32)
    got = ctx.test_bsdf(abi.NE_TEST_HG_SAMPLE, n, [u, g], 3)
    exp = np.zeros((n, 3), np.float32)
    for i in range(n):
        out = np.zeros(3, np.float32)
        O.lib().neo_hg_sample_local(1, float(g[i]), O.fp(O.f(u[i:i + 1])), O.fp(out))
        exp[i] = out
    np.testing.assert_allclose(got, exp, atol=2e-6, rtol=RTOL)
    # GGX D and pdf
    h = rng.normal(size=(n, 3))
    h[:, 2] = np.abs(h[:, 2])
    h = (h / np.linalg.norm(h, axis=1, keepdims=True)).astype(np.float32)
    alpha = rng.uniform(0.0025, 0.64, n).astype(np.float32)
    got = ctx.test_bsdf(abi.NE_TEST_GGX_D, n, [h, alpha], 1)[:, 0]
    exp = O.vec_op("neo_ggx_D", n, alpha, h, out_shape=(n,))
    assert (_rel(got, exp) <= RTOL).all()
    wo = rng.normal(size=(n, 3))
    wo[:, 2] = np.abs(wo[:, 2])
    wo = (wo / np.linalg.norm(wo, axis=1, keepdims=True)).astype(np.float32)
    got = ctx.test_bsdf(abi.NE_TEST_GGX_PDF, n, [wo, h, alpha], 1)[:, 0]
    exp = O.vec_op("neo_ggx_pdf", n, alpha, wo, h, out_shape=(n,))
    assert (_rel(got, exp) <= RTOL).all()


def test_glossy_and_forward_hg_parity(ctx):
    print_case("glossy/hg forward")
    rng = np.random.default_rng(8)
    n = 20000
    mat = np.concatenate([rng.uniform(0.1, 0.9, (n, 3)), rng.uniform(0.05, 0.8, (n, 1)), rng.uniform(0, 1, (n, 1))],
                         axis=1).astype(np.float32)
    wo = rng.normal(size=(n, 3))
    wo[:, 2] = np.abs(wo[:, 2]) + 0.01
    wo = (wo / np.linalg.norm(wo, axis=1, keepdims=True)).astype(np.float32)
    wi = rng.normal(size=(n, 3))
    wi = (wi / np.linalg.norm(wi, axis=1, keepdims=True)).astype(np.float32)
    u = rng.random((n, 3)).astype(np.float32)
    got = ctx.test_bsdf(abi.NE_TEST_GLOSSY, n, [mat, wo, wi, u], 7)
    f = np.zeros((n, 3), np.float32)
    p = np.zeros(n, np.float32)
    sm = np.zeros((n, 3), np.float32)
    O.lib().neo_glossy_c(n, O.fp(mat), O.fp(wo), O.fp(wi), O.fp(u), O.fp(f), O.fp(p), O.fp(sm))
    np.testing.assert_allclose(got[:, :3], f, rtol=RTOL, atol=1e-30)
    np.testing.assert_allclose(got[:, 3], p, rtol=RTOL, atol=1e-30)
    np.testing.assert_allclose(got[:, 4:7], sm, rtol=RTOL, atol=2e-6)
    d = wi
    g = rng.uniform(-0.95, 0.95, n).astype(np.float32)
    u2 = rng.random((n, 2)).astype(np.float32)
    got = ctx.test_bsdf(abi.NE_TEST_HG_FWD, n, [d, u2, g], 4)
    ws = O.vec_op("neo_hg_sample", n, g, d, u2, out_shape=(n, 3))
    np.testing.assert_allclose(got[:, :3], ws, rtol=RTOL, atol=2e-6)
    pp = O.vec_op("neo_hg_p", n, g, np.sum(d * got[:, :3], axis=1), out_shape=(n,))
    assert (_rel(got[:, 3], pp) <= 1e-4).all()


# ---- BVH closest hit against the oracle's brute-force scan ---------------------
@pytest.mark.parametrize("which", ["soup", "box"])
def test_closest_hit_matches_brute_force(ctx, which):
    print_case(f"closest {which}")
    s = scenes.triangle_soup() if which == "soup" else scenes.small_surface(subdiv=4)
    g = ctx.upload(s)
    w = O.World(s)
    n = 20000
    rays = scenes.random_rays(n, seed=9, origin_box=1.2) if which == "soup" else \
        scenes.rays_towards(n, target_box=0.9, origin_radius=2.5, seed=9)
    if which == "box":
        rays[:, 1] += 1.0
    t, ty, ix = ctx.test_closest(g, rays)
    bt, bty, bix = w.closest(rays, brute=True)
    np.testing.assert_array_equal(ty, bty)
    np.testing.assert_array_equal(ix, bix)
    m = bty >= 0
    assert m.mean() > 0.3
    assert (_rel(t[m], bt[m]) <= RTOL).all()


# ---- volume -------------------------------------------------------------------
def test_density_bit_exact(ctx):
    print_case("density")
    s = scenes.small_volume(res=32)
    g = ctx.upload(s)
    rng = np.random.default_rng(10)
    pts = rng.uniform(-1, 33, (20000, 3)).astype(np.float32)
    pts[:100] = np.round(pts[:100])  # exact voxel corners, incl. the upper faces
    pts[100:200, 0] = 32.0
    pts = np.clip(pts, 0, None)  # the reference aborts on negative grid coordinates
    got = ctx.test_density(g, pts)
    grid = s.volume["grid"]
    exp = np.zeros(len(pts), np.float32)
    O.lib().neo_interp_density(32, 32, 32, O.fp(grid), len(pts), O.fp(O.f(pts)), O.fp(exp))
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 7, 5), (2, 4, 4), (9, 17, 33), (6, 5, 130)])
def test_density_bit_exact_on_uneven_grids(ctx, shape):
    """The device keeps the grid in 4 x 4 x 2-voxel bricks with zero padding;
    sizes that are not multiples of the brick (and 1-voxel axes) read the
    same trilinear densities as the oracle's row-major grid, on the upper
    faces (corners at index res read 0) and at every brick boundary."""
    print_case(f"density uneven grid {shape}")
    D, H, W = shape
    rng = np.random.default_rng(sum(shape))
    grid = rng.uniform(0.0, 1.0, shape).astype(np.float32)
    s = Scene()
    s.set_volume(grid, 5.0, (0.9,) * 3, (0.1,) * 3, 0.3, maj_res=1)
    g = ctx.upload(s)
    res = np.array([W, H, D], np.float32)
    pts = rng.uniform(0.0, 1.0, (20000, 3)).astype(np.float32) * res
    # every integer corner, brick edges and the upper faces exactly
    ii = np.stack(np.meshgrid(np.arange(W + 1), np.arange(H + 1), np.arange(D + 1), indexing="ij"), -1)
    corners = ii.reshape(-1, 3).astype(np.float32)[:5000]
    near = corners[: len(corners) // 2] + np.float32(0.999)
    pts = np.concatenate([pts, corners, np.minimum(near, res)]).astype(np.float32)
    got = ctx.test_density(g, pts)
    exp = np.zeros(len(pts), np.float32)
    O.lib().neo_interp_density(W, H, D, O.fp(grid), len(pts), O.fp(O.f(pts)), O.fp(exp))
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("maj_res", [1, 4])
def test_delta_tracking_trajectories(ctx, maj_res):
    print_case(f"delta tracking maj {maj_res}")
    s = scenes.small_volume(res=32, maj_res=maj_res)
    g = ctx.upload(s)
    w = O.World(s)
    n = 20000
    rays = scenes.rays_towards(n, target_box=0.5, origin_radius=1.5, seed=11)
    sc, t = ctx.test_delta_track(g, rays, seed=3)
    osc, ot = w.delta_track(rays, seed=3)
    same = (sc == osc) & (np.abs(t - ot) <= 1e-5 * np.maximum(np.abs(ot), 1e-6))
    print(f"delta tracking identical trajectories: {same.mean():.5f}")
    assert same.mean() > 0.99
    # statistical agreement of the collision probability
    p, q = sc.mean(), osc.mean()
    assert abs(p - q) < 4 * np.sqrt(p * (1 - p) / n) + 1e-3


def test_transmittance_matches_oracle(ctx):
    print_case("transmittance")
    s = scenes.small_volume(res=32, maj_res=4)
    g = ctx.upload(s)
    w = O.World(s)
    n = 20000
    rays = scenes.rays_towards(n, target_box=0.5, origin_radius=1.5, seed=12)
    tmax = np.full(n, np.inf, np.float32)
    tr = ctx.test_transmittance(g, rays, tmax, seed=4)
    otr = w.transmittance(rays, tmax, seed=4)
    same = np.abs(tr - otr) <= 1e-4 * np.maximum(otr, 1e-6)
    print(f"transmittance identical: {same.mean():.5f}")
    assert same.mean() > 0.99
    assert abs(tr.mean() - otr.mean()) < 4 * np.sqrt((tr.var() + otr.var()) / n) + 1e-4
    # FP64 path agrees with FP32 in expectation
    tr64 = ctx.test_transmittance(g, rays, tmax, seed=4, precision=NE_FP64)
    assert abs(tr64.mean() - tr.mean()) < 4 * np.sqrt(2 * tr.var() / n) + 1e-4


def test_homogeneous_medium_matches_oracle(ctx):
    """HomogeneousMedia: closed-form transmittance (no draws) and exponential
    free flight, trajectory by trajectory against the oracle."""
    print_case("homogeneous medium tracking")
    s = scenes.homogeneous_box()
    g = ctx.upload(s)
    w = O.World(s)
    n = 20000
    rays = scenes.rays_towards(n, target_box=0.4, origin_radius=1.8, seed=15)
    rays[:, 1] += 0.1
    tmax = np.full(n, np.inf, np.float32)
    tr = ctx.test_transmittance(g, rays, tmax, seed=4)
    otr = w.transmittance(rays, tmax, seed=4)
    np.testing.assert_allclose(tr, otr, rtol=2e-5, atol=1e-7)
    assert (tr < 1).mean() > 0.5
    sc, t = ctx.test_delta_track(g, rays, seed=3)
    osc, ot = w.delta_track(rays, seed=3)
    same = (sc == osc) & (np.abs(t - ot) <= 1e-5 * np.maximum(np.abs(ot), 1e-6))
    print(f"homogeneous delta tracking identical trajectories: {same.mean():.5f}")
    assert same.mean() > 0.99


# ---- full paths -----------------------------------------------------------------
def _render_pair(ctx, s, cam, W, H, spp, bounces, seed=21, precision=NE_FP32):
    g = ctx.upload(s)
    film = ctx.film(W, H)
    ctx.render(g, cam, film, spp=spp, max_bounces=bounces, seed=seed, precision=precision)
    gpu = film.read()
    w = O.World(s, fp64=precision == NE_FP64)
    cpu = w.render(cam, W, H, spp, bounces, seed=seed, threads=8)
    return gpu, cpu


@pytest.mark.parametrize("which", ["volume", "surface", "cloud", "homogeneous", "emitters", "envmap"])
def test_render_matches_oracle(ctx, which):
    print_case(f"render {which}")
    if which == "volume":
        s, cam, bounces = scenes.small_volume(res=32, maj_res=4), scenes.VOLUME_CAMERA, 16
    elif which == "emitters":
        s, cam, bounces = scenes.emitter_mix(), scenes.BOX_CAMERA, 8
    elif which == "envmap":
        s, cam, bounces = scenes.env_map_scene(), scenes.VOLUME_CAMERA, 16
    elif which == "homogeneous":
        s, cam, bounces = scenes.homogeneous_box(), scenes.VOLUME_CAMERA, 16
    elif which == "cloud":
        s, cam, bounces = scenes.small_cloud(), scenes.VOLUME_CAMERA, 64
    else:
        s, cam, bounces = scenes.small_surface(subdiv=3), scenes.BOX_CAMERA, 8
    W = H = 32
    spp = 16
    gpu, cpu = _render_pair(ctx, s, cam, W, H, spp, bounces)
    same = np.all(np.abs(gpu - cpu) <= 1e-4 * np.maximum(np.abs(cpu), 1e-3), axis=2)
    print(f"{which}: identical pixels {same.mean():.4f}; mean gpu {gpu.mean() / spp:.5f} cpu {cpu.mean() / spp:.5f}")
    assert same.mean() >= 0.99
    # image means agree within Monte Carlo error
    assert abs(gpu.mean() - cpu.mean()) <= 0.02 * abs(cpu.mean()) + 1e-6


def test_render_fp64_mode_matches_oracle_fp64(ctx):
    print_case("render fp64")
    s = scenes.small_volume(res=32, maj_res=4)
    gpu, cpu = _render_pair(ctx, s, scenes.VOLUME_CAMERA, 24, 24, 8, 16, precision=NE_FP64)
    same = np.all(np.abs(gpu - cpu) <= 1e-9 * np.maximum(np.abs(cpu), 1e-3), axis=2)
    print(f"fp64 identical pixels {same.mean():.4f}")
    assert same.mean() >= 0.99


def test_render_is_deterministic_across_launch_shapes(ctx):
    print_case("determinism")
    s = scenes.small_volume(res=32, maj_res=4)
    g = ctx.upload(s)
    cam = scenes.VOLUME_CAMERA
    outs = []
    for spt, blk in [(0, 0), (1, 64), (3, 256), (16, 128)]:
        film = ctx.film(40, 24)
        ctx.render(g, cam, film, spp=16, max_bounces=16, seed=5, samples_per_thread=spt, block_size=blk)
        outs.append(film.read())
    for o in outs[1:]:
        np.testing.assert_allclose(o, outs[0], rtol=1e-12, atol=0)


def test_split_sample_ranges_add_up(ctx):
    print_case("sample offsets")
    s = scenes.small_surface(subdiv=2)
    g = ctx.upload(s)
    cam = scenes.BOX_CAMERA
    a = ctx.film(16, 16)
    ctx.render(g, cam, a, spp=8, max_bounces=8, seed=9)
    b = ctx.film(16, 16)
    ctx.render(g, cam, b, spp=3, max_bounces=8, seed=9)
    ctx.render(g, cam, b, spp=5, max_bounces=8, seed=9, sample_offset=3)
    np.testing.assert_allclose(a.read(), b.read(), rtol=1e-12)


def test_tile_api_matches_full_frame(ctx):
    print_case("tiles")
    from narvalengine_amd import OfflineEngine, SceneSettings
    s = scenes.small_volume(res=32, maj_res=4)
    settings = SceneSettings(resolution=(80, 40), spp=4, bounces=8)
    cam = scenes.VOLUME_CAMERA
    cam = type(cam)(position=cam.position, look_at=cam.look_at, up=cam.up, vfov=cam.vfov, aspect=2.0)
    eng = OfflineEngine(cam, settings, s, ctx=ctx, seed=3)
    for i in range(400):
        eng.renderTile(cam, i)
    full = ctx.film(80, 40)
    ctx.render(eng.gscene, cam, full, spp=4, max_bounces=8, seed=3)
    ref = full.tonemap(4)
    np.testing.assert_allclose(eng.pixels, ref, rtol=1e-6, atol=1e-7)
    # post-processing matches the oracle's OfflineEngine::postProcessing
    lin = full.read() / 4
    pp = O.vec_op("neo_post_processing", lin.size // 3, lin.reshape(-1, 3), out_shape=(lin.size // 3, 3))
    np.testing.assert_allclose(ref.reshape(-1, 3), pp, rtol=1e-5, atol=1e-6)


def test_tiles_of_an_image_below_the_tile_grid_are_empty(ctx):
    """OfflineEngine.cpp:20-21,59-60: with W < 40 the integer tile width is 0,
    renderTile's pixel loops run zero times and the pixels stay as they were."""
    print_case("empty tiles")
    from narvalengine_amd import NarvalError, OfflineEngine, SceneSettings
    s = scenes.small_volume(res=8, maj_res=2)
    eng = OfflineEngine(scenes.VOLUME_CAMERA, SceneSettings(resolution=(32, 24), spp=2, bounces=4), s, ctx=ctx)
    eng.pixels[:] = 0.25
    for i in (0, 17, 399):
        eng.renderTile(scenes.VOLUME_CAMERA, i)
    assert (eng.pixels == 0.25).all()
    assert not eng.film.read().any()
    with pytest.raises(NarvalError, match="tile index"):
        eng.renderTile(scenes.VOLUME_CAMERA, 400)
    eng.settings.spp = 0
    with pytest.raises(NarvalError, match="spp"):
        eng.renderTile(scenes.VOLUME_CAMERA, 0)


def test_invalid_arguments_are_reported(ctx):
    print_case("errors")
    from narvalengine_amd import Context, NarvalError
    from narvalengine_amd.scene import Scene
    s = Scene()
    m = s.add_glossy((0.5, 0.5, 0.5), 0.5)
    s.add_sphere((0, 0, 0), -1.0, m)
    with pytest.raises(NarvalError, match="radius"):
        ctx.upload(s)
    # non-finite geometry cannot be bounded by the BVH: rejected at upload
    for bad in ("sphere", "triangle"):
        s = Scene()
        m = s.add_glossy((0.5, 0.5, 0.5), 0.5)
        if bad == "sphere":
            s.add_sphere((0, 0, 0), float("inf"), m)
        else:
            s.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, np.inf, 0]]), np.array([[0, 1, 2]]), m)
        with pytest.raises(NarvalError, match="finite"):
            ctx.upload(s)
    g = ctx.upload(scenes.small_volume(res=8, maj_res=2))
    film = ctx.film(8, 8)
    with pytest.raises(NarvalError, match="spp"):
        ctx.render(g, scenes.VOLUME_CAMERA, film, spp=0, max_bounces=4)
    with pytest.raises(NarvalError, match="rectangle"):
        ctx.render(g, scenes.VOLUME_CAMERA, film, spp=1, max_bounces=4, rect=(0, 0, 9, 8))
    rays = scenes.rays_towards(4)
    with pytest.raises(NarvalError, match="max_bounces"):
        ctx.li(g, rays, 2048)
    with pytest.raises(NarvalError, match="max_bounces"):
        ctx.li(g, rays, -1)
    other = Context(0)
    try:
        with pytest.raises(NarvalError, match="another context"):
            other.li(g, rays, 4)
    finally:
        other.close()


def test_rejected_tile_call_leaves_the_film_unchanged(ctx):
    print_case("rejected tile call")
    from narvalengine_amd import NarvalError, OfflineEngine, SceneSettings
    s = scenes.small_volume(res=8, maj_res=2)
    eng = OfflineEngine(scenes.VOLUME_CAMERA, SceneSettings(resolution=(80, 20), spp=2, bounces=4), s, ctx=ctx)
    for i in range(400):
        eng.renderTile(scenes.VOLUME_CAMERA, i)
    before = eng.film.read()
    assert before.any()
    eng.settings.spp = 0
    for i in (0, 41, 399):
        with pytest.raises(NarvalError, match="spp"):
            eng.renderTile(scenes.VOLUME_CAMERA, i)
    np.testing.assert_array_equal(eng.film.read(), before)


def _li_case(which):
    if which == "volume":
        return scenes.small_volume(res=32, maj_res=4), scenes.VOLUME_CAMERA, 16
    if which == "emitters":
        return scenes.emitter_mix(), scenes.BOX_CAMERA, 8
    if which == "envmap":
        return scenes.env_map_scene(), scenes.VOLUME_CAMERA, 16
    if which == "homogeneous":
        return scenes.homogeneous_box(), scenes.VOLUME_CAMERA, 16
    if which == "cloud":
        return scenes.small_cloud(), scenes.VOLUME_CAMERA, 64
    return scenes.small_surface(subdiv=3), scenes.BOX_CAMERA, 8


@pytest.mark.parametrize("which", ["volume", "surface", "cloud", "homogeneous", "emitters", "envmap"])
def test_li_matches_oracle_li(ctx, which):
    """Integrator::Li (src/integrators/Integrator.h:23) through ne_li: ray i
    on stream (seed; i, sample, 0), per ray against the oracle's Li on the
    same stream.  Rays leave the camera towards random points of the scene."""
    print_case(f"ne_li {which}")
    s, cam, bounces = _li_case(which)
    g = ctx.upload(s)
    w = O.World(s)
    rng = np.random.default_rng(17)
    n = 3000
    o = np.tile(np.asarray(cam.position, np.float64), (n, 1))
    tgt = np.asarray(cam.look_at, np.float64) + rng.uniform(-0.9, 0.9, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    lit = 0
    for sample in (0, 5):
        gpu = ctx.li(g, rays, bounces, seed=29, sample=sample).astype(np.float64)
        cpu = w.li(rays, bounces, seed=29, sample=sample)
        same = np.all(np.abs(gpu - cpu) <= 1e-4 * np.maximum(np.abs(cpu), 1e-3), axis=1)
        print(f"ne_li {which} sample {sample}: identical rays {same.mean():.4f}; "
              f"mean gpu {gpu.mean():.5f} cpu {cpu.mean():.5f}")
        assert np.isfinite(gpu).all() and (gpu >= 0).all()
        assert same.mean() >= 0.99
        assert abs(gpu.mean() - cpu.mean()) <= 0.01 * abs(cpu.mean()) + 1e-6
        lit += int((cpu > 0).any(axis=1).sum())
    assert lit > n // 4  # the rays see light, so the comparison is not of zeros


def test_handles_release_in_any_order():
    print_case("release order")
    from narvalengine_amd import Context
    c = Context(0)
    g = c.upload(scenes.small_volume(res=8, maj_res=2))
    f = c.film(8, 8)
    c.render(g, scenes.VOLUME_CAMERA, f, spp=2, max_bounces=4)
    c.close()          # context handle first: scene and film keep it alive
    before = f.read()  # film still usable for reads
    assert np.isfinite(before).all()
    g.close()
    f.close()


@pytest.mark.parametrize("which", ["volume", "surface", "cloud", "homogeneous", "emitters", "envmap"])
def test_wavefront_matches_megakernel(ctx, which):
    print_case(f"wavefront vs megakernel {which}")
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    if which == "volume":
        s, cam, bounces = scenes.small_volume(res=32, maj_res=4), scenes.VOLUME_CAMERA, 16
    elif which == "emitters":
        s, cam, bounces = scenes.emitter_mix(), scenes.BOX_CAMERA, 8
    elif which == "envmap":
        s, cam, bounces = scenes.env_map_scene(), scenes.VOLUME_CAMERA, 16
    elif which == "homogeneous":
        s, cam, bounces = scenes.homogeneous_box(), scenes.VOLUME_CAMERA, 16
    elif which == "cloud":
        s, cam, bounces = scenes.small_cloud(), scenes.VOLUME_CAMERA, 64
    else:
        s, cam, bounces = scenes.small_surface(subdiv=3), scenes.BOX_CAMERA, 8
    g = ctx.upload(s)
    films = []
    for mode, pool in [(NE_MODE_MEGAKERNEL, 0), (NE_MODE_WAVEFRONT, 0), (NE_MODE_WAVEFRONT, 1000)]:
        f = ctx.film(40, 24)
        ctx.render(g, cam, f, spp=16, max_bounces=bounces, seed=13, mode=mode, pool_size=pool)
        films.append(f.read())
    # same paths, same per-path arithmetic; only the film accumulation differs
    # (FP64 sums vs 2^-30 fixed point)
    np.testing.assert_allclose(films[1], films[0], rtol=1e-6, atol=16 * 2.0 ** -30)
    np.testing.assert_array_equal(films[1], films[2])  # the pool size does not change the result


def test_wavefront_rect_with_strips_matches_megakernel(ctx):
    """A rect at an offset whose sides are not multiples of 8: the wavefront's
    path-to-pixel map (8x8 tiles, then the right strip, then the bottom strip)
    covers each pixel of the rect once per sample, as the megakernel's
    per-pixel loop does, and nothing outside it."""
    print_case("wavefront rect with strips")
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    g = ctx.upload(scenes.small_volume(res=32, maj_res=4))
    rect = (3, 5, 42, 36)  # 39 x 31 pixels: 4 x 3 full tiles, a 7-pixel strip right, 7 rows below
    films = []
    for mode, pool in [(NE_MODE_MEGAKERNEL, 0), (NE_MODE_WAVEFRONT, 0), (NE_MODE_WAVEFRONT, 512)]:
        f = ctx.film(45, 37)
        ctx.render(g, scenes.VOLUME_CAMERA, f, spp=8, max_bounces=16, seed=17, rect=rect, mode=mode,
                   pool_size=pool)
        films.append(f.read())
    np.testing.assert_allclose(films[1], films[0], rtol=1e-6, atol=8 * 2.0 ** -30)
    np.testing.assert_array_equal(films[1], films[2])
    inside = np.zeros((37, 45), bool)
    inside[5:36, 3:42] = True
    assert not films[1][~inside].any()
    assert (films[1][inside].sum(axis=-1) > 0).mean() > 0.5


def test_fused_shadow_tracking_with_a_point_light(ctx):
    """A volume-only scene lit by a point light and the sun: the shading stage
    tracks the shadow rays itself, toward a light at a finite distance (the
    ray's extent) as well as the sun. Same paths as the megakernel, and the
    pool size changes nothing."""
    print_case("fused shadow tracking, point light")
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    s = scenes.small_volume(res=32, maj_res=4, env=False)
    s.add_point_light((0.3, 0.35, -0.2), s.add_emitter((3.0, 2.5, 2.0)))
    g = ctx.upload(s)
    films = []
    for mode, pool in [(NE_MODE_MEGAKERNEL, 0), (NE_MODE_WAVEFRONT, 0), (NE_MODE_WAVEFRONT, 1000)]:
        f = ctx.film(40, 24)
        ctx.render(g, scenes.VOLUME_CAMERA, f, spp=16, max_bounces=16, seed=29, mode=mode, pool_size=pool)
        films.append(f.read())
    assert films[1].any()
    np.testing.assert_allclose(films[1], films[0], rtol=1e-6, atol=16 * 2.0 ** -30)
    np.testing.assert_array_equal(films[1], films[2])


@pytest.mark.parametrize("which", ["cloud", "point_light"])
def test_wide_grid_offsets_change_nothing(ctx, which):
    """NE_FLAG_WIDE_OFFSETS runs the kernels a grid of 4 GiB or more takes
    (64-bit density offsets: medium, fused shading, shadow stage) on a small
    grid. They read the same voxels, so the films and event counts equal the
    32-bit-offset kernels', plain and counted, in a full pool and a draining
    one."""
    print_case(f"wide grid offsets {which}")
    from narvalengine_amd.abi import NE_MODE_WAVEFRONT
    if which == "cloud":
        s = scenes.small_cloud()
    else:
        s = scenes.small_volume(res=32, maj_res=4, env=False)
        s.add_point_light((0.3, 0.35, -0.2), s.add_emitter((3.0, 2.5, 2.0)))
    g = ctx.upload(s)
    for pool, count in [(0, False), (1000, False), (0, True)]:
        out = []
        for wide in (False, True):
            f = ctx.film(40, 24)
            ctx.render(g, scenes.VOLUME_CAMERA, f, spp=8, max_bounces=32, seed=41, mode=NE_MODE_WAVEFRONT,
                       pool_size=pool, count=count, wide=wide)
            out.append((f.read(), ctx.counters() if count else None))
        assert out[0][0].any()
        np.testing.assert_array_equal(out[0][0], out[1][0])
        assert out[0][1] == out[1][1]


def test_plan_over_many_blocks_matches_one_block(ctx):
    """The plan scans the chunks' free slots in blocks of 1024 chunks chained
    by look-back: a 2^21-slot pool (8192 chunks of 256 slots, 8 plan blocks)
    renders the same film as a 1024-slot pool (4 chunks, one plan block)."""
    print_case("plan over many blocks")
    from narvalengine_amd.abi import NE_MODE_WAVEFRONT
    g = ctx.upload(scenes.small_volume(res=32, maj_res=4))
    films = []
    for pool in (1 << 21, 1024):
        f = ctx.film(256, 256)
        ctx.render(g, scenes.VOLUME_CAMERA, f, spp=32, max_bounces=16, seed=23, mode=NE_MODE_WAVEFRONT,
                   pool_size=pool)
        films.append(f.read())
    assert films[0].any()
    np.testing.assert_array_equal(films[0], films[1])


def test_plan_look_back_that_gives_up_fails_the_render(ctx):
    """The failure path of the plan's bounded look-back wait, forced from the
    host (NE_FLAG_TEST_PLAN_STALL: the plan blocks wait for words of a later
    launch, which never come, for 1024 spins): every kernel still ends, the
    render fails with NE_ERR_HIP instead of hanging, and the context then
    renders the same film as before."""
    print_case("plan look-back gives up")
    from narvalengine_amd import NarvalError
    from narvalengine_amd.abi import NE_MODE_WAVEFRONT
    g = ctx.upload(scenes.small_volume(res=32, maj_res=4))
    kw = dict(spp=32, max_bounces=16, seed=23, mode=NE_MODE_WAVEFRONT, pool_size=1 << 21)
    f0 = ctx.film(256, 256)
    ctx.render(g, scenes.VOLUME_CAMERA, f0, **kw)
    with pytest.raises(NarvalError, match="launch_wavefront"):
        ctx.render(g, scenes.VOLUME_CAMERA, ctx.film(256, 256), plan_stall=True, **kw)
    f1 = ctx.film(256, 256)
    ctx.render(g, scenes.VOLUME_CAMERA, f1, **kw)
    assert f0.read().any()
    np.testing.assert_array_equal(f0.read(), f1.read())


@pytest.mark.parametrize("which", ["volume", "surface"])
def test_wavefront_counters_and_stage_times(ctx, which):
    """Both pipelines trace the same paths: equal event counts, except that the
    wavefront's speculative traversal (parked leaves) visits a superset of the
    megakernel's BVH nodes and primitives; the per-launch stage times add up
    to the whole-render time."""
    print_case(f"wavefront counters/stage times {which}")
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    if which == "volume":
        s, cam, bounces = scenes.small_volume(res=32, maj_res=4), scenes.VOLUME_CAMERA, 16
    else:
        s, cam, bounces = scenes.small_surface(subdiv=3), scenes.BOX_CAMERA, 8
    g = ctx.upload(s)
    counts = {}
    for mode in (NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT):
        f = ctx.film(40, 24)
        ctx.render(g, cam, f, spp=16, max_bounces=bounces, seed=5, mode=mode, count=True)
        counts[mode] = ctx.counters()
    mk, wf = counts[NE_MODE_MEGAKERNEL], counts[NE_MODE_WAVEFRONT]
    for k in ("paths", "ext_rays", "shadow_rays", "density_lookups", "majorant_cells"):
        assert mk[k] == wf[k], (k, mk[k], wf[k])
    for k in ("bvh_nodes", "prim_tests"):
        assert mk[k] <= wf[k] <= 1.5 * mk[k], (k, mk[k], wf[k])
    assert wf["paths"] == 40 * 24 * 16
    assert all(mk["shadow_" + k] == 0 for k in ("density_lookups", "majorant_cells", "bvh_nodes", "prim_tests"))
    if which == "volume":
        assert 0 < wf["shadow_density_lookups"] < wf["density_lookups"]
    else:
        assert 0 < wf["shadow_bvh_nodes"] < wf["bvh_nodes"]
    f = ctx.film(40, 24)
    ctx.render(g, cam, f, spp=16, max_bounces=bounces, seed=5, mode=NE_MODE_WAVEFRONT, timed=True,
               stage_timed=True)
    total = ctx.kernel_ms()
    st = ctx.stage_times()
    assert st["megakernel"] == (0.0, 0)
    n = {k: v[1] for k, v in st.items()}
    # the plan (and the regen stage) stop once every path has been issued;
    # the last iteration plans and ends
    assert 1 <= n["plan"] <= n["shade"] + 1 and n["shade"] >= 1
    # volume-only plain renders track their shadow rays in the shading stage
    assert n["shadow"] == (0 if which == "volume" else n["shade"])
    # a volume-only scene makes its new paths inside the medium stage
    if which == "volume":
        assert n["regen"] == 0
    else:
        assert 1 <= n["regen"] <= n["plan"]
    assert (n["trace"] == n["shade"]) == (which == "surface")
    assert (n["medium"] == n["shade"]) == (which == "volume")
    staged = sum(v[0] for v in st.values())
    assert 0 < staged <= total * 1.02 + 0.05 and staged >= 0.5 * total


@pytest.mark.parametrize("which", ["surface", "cloud", "emitters"])
def test_material_sort_leaves_the_image_unchanged(ctx, which):
    """Shading each chunk in (kind, material) order is a permutation of the
    same paths: the fixed-point film and the event counts are identical (BVH
    node and primitive counts aside: speculative traversal visits depend on
    which rays share a wave)."""
    print_case(f"material sort {which}")
    from narvalengine_amd.abi import NE_MODE_WAVEFRONT
    if which == "cloud":
        s, cam, bounces = scenes.small_cloud(), scenes.VOLUME_CAMERA, 64
    elif which == "emitters":
        s, cam, bounces = scenes.emitter_mix(), scenes.BOX_CAMERA, 8
    else:
        s, cam, bounces = scenes.small_surface(subdiv=3), scenes.BOX_CAMERA, 8
    g = ctx.upload(s)
    out = []
    for srt in (False, True):
        f = ctx.film(40, 24)
        ctx.render(g, cam, f, spp=16, max_bounces=bounces, seed=17, mode=NE_MODE_WAVEFRONT, sort=srt, count=True)
        out.append((f.read(), ctx.counters()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    wave_dependent = ("bvh_nodes", "prim_tests", "shadow_bvh_nodes", "shadow_prim_tests")
    c0, c1 = out[0][1], out[1][1]
    assert {k: v for k, v in c0.items() if k not in wave_dependent} == \
        {k: v for k, v in c1.items() if k not in wave_dependent}


def test_empty_space_skipping_is_exact(ctx):
    """Empty majorant cells hold -D, D the taxicab distance to the nearest
    non-empty cell, and the DDA steps over the next D - 1 cells without reading
    them: non-empty majorants equal the oracle's, and images and event counts
    are identical with the skip on and off, megakernel and wavefront."""
    print_case("empty-space skipping")
    from scipy.ndimage import distance_transform_cdt

    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    M = 12
    on = scenes.small_cloud(res=48, maj_res=M)
    off = scenes.small_cloud(res=48, maj_res=M)
    off.volume["empty_skip"] = False
    g_on, g_off = ctx.upload(on), ctx.upload(off)
    m_on, m_off = ctx.test_majorants(g_on, M), ctx.test_majorants(g_off, M)
    om = O.World(on).majorants(M).reshape(M, M, M)
    np.testing.assert_array_equal(m_off, om)
    empty = om <= 0
    assert 0.2 < empty.mean() < 0.95
    dist = distance_transform_cdt(empty, metric="taxicab")
    assert dist.max() >= 3  # runs long enough to skip
    np.testing.assert_array_equal(m_on[~empty], om[~empty])
    np.testing.assert_array_equal(m_on[empty], -dist[empty].astype(np.float32))
    for mode in (NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT):
        out = []
        for g in (g_on, g_off):
            f = ctx.film(40, 24)
            ctx.render(g, scenes.VOLUME_CAMERA, f, spp=16, max_bounces=64, seed=21, mode=mode, count=True)
            out.append((f.read(), ctx.counters()))
            f.close()
        np.testing.assert_array_equal(out[0][0], out[1][0])
        assert out[0][1] == out[1][1]
        assert out[0][1]["majorant_cells"] > out[0][1]["density_lookups"] > 0
    g_on.close()
    g_off.close()


def test_wavefront_pool_rounds(ctx):
    """Pools above 16384 chunks take several rounds of the planning scan
    (here 32771 chunks: two full rounds and a ragged third): the image equals
    the one from a single-round pool and from a 196-chunk pool."""
    print_case("wavefront pool rounds")
    from narvalengine_amd.abi import NE_MODE_WAVEFRONT
    s = scenes.small_volume(res=32, maj_res=4)
    g = ctx.upload(s)
    films = []
    for pool in (1 << 22, (1 << 23) + 3 * 256, 50000):
        f = ctx.film(64, 64)
        ctx.render(g, scenes.VOLUME_CAMERA, f, spp=2100, max_bounces=4, seed=23, mode=NE_MODE_WAVEFRONT,
                   pool_size=pool, count=pool == 50000)
        films.append(f.read())
        f.close()
    assert ctx.counters()["paths"] == 64 * 64 * 2100
    np.testing.assert_array_equal(films[0], films[1])
    np.testing.assert_array_equal(films[0], films[2])
    g.close()


def test_tracking_on_a_sparse_grid_matches_oracle(ctx):
    """A dense block in an empty grid (clipped walk, empty-cell skipping):
    device delta-tracking trajectories and ratio-tracking transmittances
    agree with the oracle's, ray by ray."""
    print_case("sparse-grid tracking")
    import test_oracle_invariants as inv
    s, _ = inv._sparse_volume()
    g = ctx.upload(s)
    w = O.World(s)
    n = 20000
    rays = scenes.rays_towards(n, target_box=0.4, origin_radius=1.2, seed=31)
    sc, t = ctx.test_delta_track(g, rays, seed=7)
    osc, ot = w.delta_track(rays, seed=7)
    same = (sc == osc) & (np.abs(t - ot) <= 1e-5 * np.maximum(np.abs(ot), 1e-6))
    print(f"sparse delta tracking identical: {same.mean():.5f}, collisions {sc.mean():.3f}")
    assert same.mean() > 0.99 and 0.05 < sc.mean() < 0.95
    tmax = np.full(n, np.inf, np.float32)
    tr = ctx.test_transmittance(g, rays, tmax, seed=8)
    otr = w.transmittance(rays, tmax, seed=8)
    same = np.abs(tr - otr) <= 1e-4 * np.maximum(otr, 1e-6)
    print(f"sparse transmittance identical: {same.mean():.5f}")
    assert same.mean() > 0.99
    g.close()


@pytest.mark.parametrize("which", ["volume", "surface", "emitters"])
def test_per_sample_film_matches_fixed_point_film(ctx, which):
    """The wavefront's two film paths: per-sample radiance slots summed per
    pixel in FP64 (the default while the slots fit) and 2^-30 fixed-point
    atomics (larger renders) give the same image up to the fixed-point
    rounding, and each is reproducible."""
    print_case(f"per-sample vs fixed-point film {which}")
    from narvalengine_amd.abi import NE_MODE_WAVEFRONT
    if which == "volume":
        s, cam, bounces = scenes.small_volume(res=32, maj_res=4), scenes.VOLUME_CAMERA, 16
    elif which == "emitters":
        s, cam, bounces = scenes.emitter_mix(), scenes.BOX_CAMERA, 8
    else:
        s, cam, bounces = scenes.small_surface(subdiv=3), scenes.BOX_CAMERA, 8
    g = ctx.upload(s)
    out = {}
    for fixed in (False, True, False):
        f = ctx.film(40, 24)
        ctx.render(g, cam, f, spp=32, max_bounces=bounces, seed=19, mode=NE_MODE_WAVEFRONT, fixed_film=fixed,
                   pool_size=3000)
        out.setdefault(fixed, []).append(f.read())
        f.close()
    np.testing.assert_array_equal(out[False][0], out[False][1])
    # up to 32 samples x a few contributions each, rounded to 2^-30
    np.testing.assert_allclose(out[True][0], out[False][0], rtol=1e-6, atol=256 * 2.0 ** -30)
    assert out[False][0].sum() > 0
    g.close()
