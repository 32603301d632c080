"""Batched rendering, both executors (persistent kernel, wavefront lanes):
per-job sums bitwise independent of the batch and of the work-item shape or
lane order, equal to the single-job render up to the film's fixed-point
quantisation, and statistically equal to the oracle."""
import numpy as np
import pytest

import oracle as O
import scenes
from conftest import print_case
from narvalengine_amd import Batch, configs
from narvalengine_amd.abi import NE_BATCH_WAVEFRONT, NE_FP32, NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
from narvalengine_amd.scene import Camera

pytestmark = pytest.mark.gpu


def _jobs(ctx):
    specs = [(16, 1, 40, 24, 8), (24, 2, 33, 17, 5), (16, 3, 20, 40, 12), (32, 4, 8, 8, 3)]
    out = []
    for res, seed, w, h, spp in specs:
        s = scenes.small_volume(res=res, maj_res=4)
        s.volume["grid"] = configs.fbm_grid(res, seed)
        cam = Camera(position=(0.0, 0.25, -1.8), look_at=(0, 0, 0), up=(0, 1, 0), vfov=45.0, aspect=w / h)
        out.append(dict(scene=ctx.upload(s), scene_host=s, camera=cam, width=w, height=h, spp=spp, max_bounces=12,
                        seed=7 + seed, job=seed))
    return out


def test_batch_is_bitwise_independent_of_batching_and_item_shape(ctx):
    print_case("batch vs single-job batches, two item shapes")
    jobs = _jobs(ctx)
    b = Batch(ctx, jobs)
    ms = b.run((8, 8, 4))
    assert ms > 0
    full = [b.read(k) for k in range(len(jobs))]
    b.run((16, 4, 1))
    for k in range(len(jobs)):
        np.testing.assert_array_equal(b.read(k), full[k])
    for k, j in enumerate(jobs):
        single = Batch(ctx, [j])
        single.run((32, 8, 16))
        np.testing.assert_array_equal(single.read(0), full[k])
        single.close()
    t = b.job_times()
    assert (t > 0).all() and (t <= b.run((16, 4, 1)) * 1.05 + 0.5).all()
    b.close()


def test_batch_matches_the_single_job_render(ctx):
    """Same paths as ne_render (same counter RNG streams, same per-path
    arithmetic); only the accumulation differs (FP64 sums vs 2^-30 fixed point)."""
    print_case("batch vs megakernel render")
    jobs = _jobs(ctx)
    b = Batch(ctx, jobs)
    b.run()
    for k, j in enumerate(jobs):
        f = ctx.film(j["width"], j["height"])
        ctx.render(j["scene"], j["camera"], f, spp=j["spp"], max_bounces=j["max_bounces"], seed=j["seed"],
                   job=j["job"], mode=NE_MODE_MEGAKERNEL)
        ref = f.read()
        np.testing.assert_allclose(b.read(k), ref, rtol=1e-6, atol=j["spp"] * 2.0 ** -30)
    b.close()


def test_wavefront_executor_is_bitwise_independent_of_batching(ctx):
    """The wavefront executor runs the jobs on several streams, longest first:
    which lane takes a job, and what runs beside it, leaves its sums bitwise
    unchanged (fixed-point film)."""
    print_case("batch wavefront executor vs single-job batches")
    jobs = _jobs(ctx)
    wf = (8, 8, 4, NE_BATCH_WAVEFRONT)
    b = Batch(ctx, jobs)
    assert b.run(wf) > 0
    full = [b.read(k) for k in range(len(jobs))]
    b.run(wf)
    for k in range(len(jobs)):
        np.testing.assert_array_equal(b.read(k), full[k])
    for k, j in enumerate(jobs):
        single = Batch(ctx, [j])
        single.run(wf)
        np.testing.assert_array_equal(single.read(0), full[k])
        single.close()
    t = b.job_times()
    assert (t > 0).all()
    b.close()


def test_wavefront_executor_matches_the_wavefront_render(ctx):
    """Same pipeline and paths as ne_render's wavefront; only the accumulation
    differs (2^-30 fixed point for every path against FP64 per-sample sums)."""
    print_case("batch wavefront executor vs wavefront render")
    jobs = _jobs(ctx)
    b = Batch(ctx, jobs)
    b.run((8, 8, 4, NE_BATCH_WAVEFRONT))
    for k, j in enumerate(jobs):
        f = ctx.film(j["width"], j["height"])
        ctx.render(j["scene"], j["camera"], f, spp=j["spp"], max_bounces=j["max_bounces"], seed=j["seed"],
                   job=j["job"], mode=NE_MODE_WAVEFRONT)
        np.testing.assert_allclose(b.read(k), f.read(), rtol=1e-6, atol=4 * j["spp"] * 2.0 ** -30)
    b.close()


def test_both_executors_on_mixed_scenes(ctx):
    """Surface, emitter, homogeneous-medium and grid-volume jobs in one batch:
    each executor matches its single-job render (the persistent kernel the
    megakernel, the wavefront executor the wavefront) up to the fixed-point
    film, and the two executors agree within the same bound."""
    print_case("batch executors on mixed scenes")
    specs = [(scenes.small_surface(subdiv=2), scenes.BOX_CAMERA, 24, 24, 6, 6),
             (scenes.emitter_mix(), scenes.BOX_CAMERA, 20, 16, 5, 6),
             (scenes.homogeneous_box(), scenes.BOX_CAMERA, 16, 20, 4, 8),
             (scenes.small_volume(res=16, maj_res=2), scenes.VOLUME_CAMERA, 18, 14, 7, 12)]
    jobs = [dict(scene=ctx.upload(s), camera=cam, width=w, height=h, spp=spp, max_bounces=mb, seed=40 + k, job=k)
            for k, (s, cam, w, h, spp, mb) in enumerate(specs)]
    per = Batch(ctx, jobs)
    per.run((8, 8, 2))
    wf = Batch(ctx, jobs)
    wf.run((8, 8, 2, NE_BATCH_WAVEFRONT))
    for k, j in enumerate(jobs):
        tol = dict(rtol=1e-6, atol=4 * j["spp"] * 2.0 ** -30)
        for mode, b in ((NE_MODE_MEGAKERNEL, per), (NE_MODE_WAVEFRONT, wf)):
            f = ctx.film(j["width"], j["height"])
            ctx.render(j["scene"], j["camera"], f, spp=j["spp"], max_bounces=j["max_bounces"], seed=j["seed"],
                       job=j["job"], mode=mode)
            np.testing.assert_allclose(b.read(k), f.read(), **tol)
        np.testing.assert_allclose(per.read(k), wf.read(k), rtol=2e-6, atol=8 * j["spp"] * 2.0 ** -30)
        assert per.read(k).sum() > 0
    per.close()
    wf.close()


def test_batch_matches_the_oracle(ctx):
    print_case("batch vs oracle")
    jobs = _jobs(ctx)[:2]
    b = Batch(ctx, jobs)
    b.run()
    for k, j in enumerate(jobs):
        w = O.World(j["scene_host"])
        ref = w.render(j["camera"], j["width"], j["height"], j["spp"], j["max_bounces"], seed=j["seed"], job=j["job"])
        got = b.read(k)
        same = np.isclose(got, ref, rtol=1e-5, atol=1e-6).mean()
        assert same >= 0.99, same
        assert abs(got.mean() - ref.mean()) < 0.02 * ref.mean()
    b.close()


def test_batch_autotune_picks_a_candidate_and_reports_times(ctx):
    print_case("batch autotune")
    jobs = _jobs(ctx)
    b = Batch(ctx, jobs)
    cfg, ms = b.autotune(spp_divisor=2)
    assert ms > 0 and cfg[0] * cfg[1] >= 32 and cfg[2] >= 1
    full = b.run()  # tuned shape
    assert full > 0 and len(b.job_times()) == len(jobs)
    b.close()


def test_batch_rejects_bad_jobs(ctx):
    print_case("batch invalid arguments")
    from narvalengine_amd.abi import NarvalError
    jobs = _jobs(ctx)[:1]
    bad = dict(jobs[0], spp=0)
    with pytest.raises(NarvalError):
        Batch(ctx, [bad])
    b = Batch(ctx, jobs)
    with pytest.raises(NarvalError):
        b.run((0, 8, 4))
    with pytest.raises(NarvalError):
        b.job_times()
    b.close()
