"""Parity at the benchmark workloads' full sizes, through properties that do
not depend on size: the 1,310,720-triangle surface1m scene and the 512^3
cloud512 grid, uploaded exactly as the benchmark uploads them.

* surface1m: BVH closest hits against the oracle's BVH (20,000 rays) and its
  brute-force scan (64 rays): hit type and original index bit-exact, t within
  1e-5 relative.  Megakernel and wavefront renders of the full 1024x1024 frame
  trace the same paths (equal to 1e-6, the film rounding).
* cloud512: the 16^3 majorant grid equals the oracle's, empty cells carry
  minus their taxicab distance to the nearest non-empty cell; delta-tracking
  trajectories and ratio-tracking transmittances match the oracle ray by ray;
  megakernel and wavefront renders of the cloud trace the same paths.
* batch48: the 48 uneven jobs in one persistent launch; the three smallest
  jobs are bitwise equal to their own single-job batches, and one equals the
  single-job megakernel render up to the film rounding.
"""
import numpy as np
import pytest

import oracle as O
import scenes
from conftest import print_case
from narvalengine_amd import configs

pytestmark = pytest.mark.gpu
RTOL = 1e-5


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


@pytest.fixture(scope="module")
def surface(ctx):
    wl = configs.surface1m()
    g = ctx.upload(wl.scene)
    yield wl, g, O.World(wl.scene)
    g.close()


@pytest.fixture(scope="module")
def cloud(ctx):
    wl = configs.cloud512()
    g = ctx.upload(wl.scene)
    yield wl, g, O.World(wl.scene)
    g.close()


def test_full_surface_closest_hits(ctx, surface):
    print_case("surface1m closest hits")
    wl, g, w = surface
    assert sum(len(v) for v in wl.scene.tri_v) == 1310720
    assert g.info()["bvh_nodes"] > 1310720 // 32
    box = scenes.rays_towards(10000, target_box=0.9, origin_radius=2.5, seed=61)
    box[:, 1] += 1.0
    # rays aimed at the displaced icosphere (radius 0.45 about (0, 0.55, 0.35))
    ico = scenes.rays_towards(10000, target_box=0.4, origin_radius=1.2, seed=63)
    ico[:, :3] += np.array([0.0, 0.55, 0.35], np.float32)
    rays = np.concatenate([ico, box])
    t, ty, ix = ctx.test_closest(g, rays)
    ot, oty, oix = w.closest(rays)
    np.testing.assert_array_equal(ty, oty)
    np.testing.assert_array_equal(ix, oix)
    m = oty >= 0
    assert m.mean() > 0.9
    assert (_rel(t[m], ot[m]) <= RTOL).all()
    assert (oty[:10000] == 0).mean() > 0.2  # icosphere triangles (the walls block the rest)
    bt, bty, bix = w.closest(rays[:64], brute=True)
    np.testing.assert_array_equal(ty[:64], bty)
    np.testing.assert_array_equal(ix[:64], bix)


def test_full_surface_frame_same_paths_in_both_pipelines(ctx, surface):
    print_case("surface1m full frame megakernel vs wavefront")
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    wl, g, _ = surface
    films = []
    for mode in (NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT):
        f = ctx.film(wl.width, wl.height)
        ctx.render(g, wl.camera, f, spp=1, max_bounces=wl.max_bounces, seed=17, mode=mode)
        films.append(f.read())
        f.close()
    assert np.isfinite(films[0]).all() and (films[0] >= 0).all()
    assert films[0].mean() > 0.01
    np.testing.assert_allclose(films[1], films[0], rtol=1e-6, atol=16 * 2.0 ** -30)


def test_full_cloud_majorants_and_skip_distances(ctx, cloud):
    print_case("cloud512 majorants")
    from scipy.ndimage import distance_transform_cdt
    wl, g, w = cloud
    M = 16
    m = ctx.test_majorants(g, M)
    om = w.majorants(M).reshape(M, M, M)
    empty = om <= 0
    assert 0.05 < empty.mean() < 0.95
    np.testing.assert_array_equal(m[~empty], om[~empty])
    dist = distance_transform_cdt(empty, metric="taxicab")
    np.testing.assert_array_equal(m[empty], -dist[empty].astype(np.float32))


def test_full_cloud_tracking_matches_oracle(ctx, cloud):
    print_case("cloud512 tracking")
    wl, g, w = cloud
    n = 4000
    rays = scenes.rays_towards(n, target_box=0.4, origin_radius=1.5, seed=62)
    sc, t = ctx.test_delta_track(g, rays, seed=9)
    osc, ot = w.delta_track(rays, seed=9)
    same = (sc == osc) & (np.abs(t - ot) <= 1e-5 * np.maximum(np.abs(ot), 1e-6))
    print(f"cloud512 delta tracking identical: {same.mean():.5f}, collisions {sc.mean():.3f}")
    assert same.mean() > 0.99 and sc.mean() > 0.05
    tmax = np.full(n, np.inf, np.float32)
    tr = ctx.test_transmittance(g, rays, tmax, seed=10)
    otr = w.transmittance(rays, tmax, seed=10)
    same = np.abs(tr - otr) <= 1e-4 * np.maximum(otr, 1e-6)
    print(f"cloud512 transmittance identical: {same.mean():.5f}")
    assert same.mean() > 0.99


def test_full_cloud_same_paths_in_both_pipelines(ctx, cloud):
    print_case("cloud512 megakernel vs wavefront")
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    wl, g, _ = cloud
    W, H = 192, 108
    cam = type(wl.camera)(position=wl.camera.position, look_at=wl.camera.look_at, up=wl.camera.up,
                          vfov=wl.camera.vfov, aspect=W / H)
    films = []
    for mode in (NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT):
        f = ctx.film(W, H)
        ctx.render(g, cam, f, spp=4, max_bounces=wl.max_bounces, seed=23, mode=mode)
        films.append(f.read())
        f.close()
    assert np.isfinite(films[0]).all() and (films[0] >= 0).all()
    assert films[0].mean() > 0.01
    np.testing.assert_allclose(films[1], films[0], rtol=1e-6, atol=64 * 2.0 ** -30)


def test_full_batch48_jobs_are_independent_of_the_batch(ctx):
    print_case("batch48 full batch vs single jobs")
    from narvalengine_amd import Batch
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL
    wls = configs.batch48()
    jobs = []
    for k, w in enumerate(wls):
        jobs.append(dict(scene=ctx.upload(w.scene), camera=w.camera, width=w.width, height=w.height, spp=w.spp,
                         max_bounces=w.max_bounces, seed=1, job=k))
    assert len(jobs) == 48
    b = Batch(ctx, jobs)
    b.run()
    sums = [b.read(k) for k in range(48)]
    for k, j in enumerate(jobs):
        assert np.isfinite(sums[k]).all() and sums[k].mean() > 0, k
    small = sorted(range(48), key=lambda k: jobs[k]["width"] * jobs[k]["height"] * jobs[k]["spp"])[:3]
    for k in small:
        one = Batch(ctx, [jobs[k]])
        one.run()
        np.testing.assert_array_equal(one.read(0), sums[k])
        one.close()
    j = jobs[small[0]]
    f = ctx.film(j["width"], j["height"])
    ctx.render(j["scene"], j["camera"], f, spp=j["spp"], max_bounces=j["max_bounces"], seed=j["seed"], job=j["job"],
               mode=NE_MODE_MEGAKERNEL)
    np.testing.assert_allclose(sums[small[0]], f.read(), rtol=1e-6, atol=j["spp"] * 2.0 ** -30)
    f.close()
    b.close()
    for j in jobs:
        j["scene"].close()
