"""Learned guide on the device against its CPU restatement
(oracle/guide_ref.py): MFMA forward, mixture head, loss and gradients, and a
training run that must fit a known directional distribution."""
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import print_case
from narvalengine_amd import Guide

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import guide_ref as G  # noqa: E402

pytestmark = pytest.mark.gpu
LOG2T = 12


def _random_params(g, seed):
    rng = np.random.default_rng(seed)
    table = rng.normal(0, 0.5, g.table_floats).astype(np.float32)
    mlp = np.zeros(g.mlp_floats, np.float32)
    for l in range(G.LAYERS):
        used_in = 35 if l == 0 else 64
        used_out = 40 if l == G.LAYERS - 1 else 64
        w = np.zeros((G.IN[l], G.OUT[l]), np.float32)
        w[:used_in, :used_out] = rng.normal(0, np.sqrt(2 / used_in), (used_in, used_out))
        mlp[G.W_OFF[l]:G.W_OFF[l + 1]] = w.reshape(-1)
        mlp[G.B_OFF[l]:G.B_OFF[l] + used_out] = rng.normal(0, 0.1, used_out)
    return table, mlp


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([p, d], 1).astype(np.float32)


def test_guide_forward_matches_reference(ctx):
    print_case("guide forward (MFMA) vs FP64 reference")
    g = Guide(ctx, log2_table=LOG2T, batch=4096, seed=1)
    table, mlp = _random_params(g, 2)
    g.set_params(table, mlp)
    q = _queries(777, 3)  # not a multiple of 16: ragged last tile
    raw = g.eval(q, raw=True)
    ref, _ = G.forward(q, table, mlp, g.level_res, LOG2T, quantize=True)
    err = np.abs(raw - ref)
    scale = np.abs(ref).mean()
    assert np.median(err) < 2e-3 * scale + 1e-4, (np.median(err), scale)
    assert err.max() < 0.1 * scale + 1e-2, (err.max(), scale)
    # the mixture head on the device equals the closed form applied to the raw outputs
    mix = g.eval(q).reshape(-1, 8, 5)
    pi, mu, kap = G.head(raw)
    np.testing.assert_allclose(mix[:, :, 0], pi, rtol=2e-5, atol=1e-7)
    np.testing.assert_allclose(mix[:, :, 1:4], mu, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(mix[:, :, 4], kap, rtol=2e-5)
    np.testing.assert_allclose(mix[:, :, 0].sum(1), 1.0, rtol=1e-5)
    g.close()


def test_guide_gradients_match_reference(ctx):
    print_case("guide gradients vs FP64 reference")
    g = Guide(ctx, log2_table=LOG2T, batch=4096, seed=1)
    table, mlp = _random_params(g, 5)
    g.set_params(table, mlp)
    rng = np.random.default_rng(6)
    q = _queries(1000, 7)
    w = rng.normal(size=(1000, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    r = rng.uniform(0.1, 2.0, 1000)
    s = np.concatenate([q, w, r[:, None]], 1).astype(np.float32)
    loss = g.train_step(s, update=False)
    gt, gm = g.grads()
    rl, rt, rm = G.loss_and_grad(s, table, mlp, g.level_res, LOG2T, quantize=True)
    assert abs(loss - rl) < 1e-3 * abs(rl) + 1e-4, (loss, rl)
    for got, ref in ((gm, rm), (gt, rt)):
        cos = got @ ref / (np.linalg.norm(got) * np.linalg.norm(ref))
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert cos > 0.995 and rel < 0.08, (cos, rel)
    # padding entries of the weight matrices get no gradient
    for l in range(G.LAYERS):
        W = gm[G.W_OFF[l]:G.W_OFF[l + 1]].reshape(G.IN[l], G.OUT[l])
        if l == 0:
            assert not W[35:].any()
        if l == G.LAYERS - 1:
            assert not W[:, 40:].any()
    g.close()


def test_guide_learns_a_position_dependent_lobe(ctx):
    """Samples drawn from one vMF lobe whose mean depends on position: training
    must raise the proposal density of the true mean direction."""
    print_case("guide training")
    g = Guide(ctx, log2_table=14, batch=8192, seed=3, lr_table=1e-2, lr_mlp=2e-3)
    rng = np.random.default_rng(8)

    def target_mu(p):
        m = np.stack([np.cos(3 * p[:, 0]), np.sin(3 * p[:, 0]), 2 * p[:, 1] - 1], 1)
        return m / np.linalg.norm(m, axis=1, keepdims=True)

    def vmf_sample(mu, kappa, n):
        u1, u2 = rng.uniform(size=n), rng.uniform(size=n)
        c = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
        s = np.sqrt(np.maximum(0, 1 - c * c))
        ph = 2 * np.pi * u2
        a = np.where(np.abs(mu[:, 2:3]) < 0.9, np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]))
        t1 = np.cross(mu, a)
        t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
        t2 = np.cross(mu, t1)
        return c[:, None] * mu + (s * np.cos(ph))[:, None] * t1 + (s * np.sin(ph))[:, None] * t2

    test_q = _queries(512, 9)
    mu_t = target_mu(test_q.astype(np.float64))

    def density_at_target():
        mix = g.eval(test_q)
        y = np.zeros((len(test_q), 40))
        y[:, 0::5] = np.log(np.maximum(mix[:, 0::5], 1e-30))
        y[:, 1::5], y[:, 2::5], y[:, 3::5] = mix[:, 1::5], mix[:, 2::5], mix[:, 3::5]
        y[:, 4::5] = np.log(mix[:, 4::5])
        return G.mixture_pdf(y, mu_t).mean()

    before = density_at_target()
    losses = []
    for step in range(150):
        q = _queries(8192, 100 + step)
        mu = target_mu(q.astype(np.float64))
        w = vmf_sample(mu, 20.0, len(q))
        s = np.concatenate([q, w, np.ones((len(q), 1))], 1).astype(np.float32)
        losses.append(g.train_step(s))
    after = density_at_target()
    assert np.mean(losses[-10:]) < np.mean(losses[:10]) - 1.0, (losses[:3], losses[-3:])
    assert after > 5 * before and after > 1.0, (before, after)
    st = g.stats()
    assert st["steps"] == 150 and st["samples"] == 150 * 8192
    g.close()


def test_guide_rejects_bad_arguments(ctx):
    print_case("guide invalid arguments")
    from narvalengine_amd.abi import NarvalError
    with pytest.raises(NarvalError):
        Guide(ctx, log2_table=30)
    g = Guide(ctx, log2_table=LOG2T, batch=64)
    bad = np.zeros((4, 10), np.float32)
    bad[0, 9] = -1.0
    with pytest.raises(NarvalError):
        g.train_step(bad)
    with pytest.raises(NarvalError):
        g.train_step(np.zeros((65, 10), np.float32))
    g.close()


def _cloud():
    import scenes
    return scenes.small_cloud(res=48, maj_res=6), scenes.VOLUME_CAMERA


@pytest.mark.parametrize("trained", [False, True])
def test_guided_render_is_unbiased(ctx, trained):
    """One-sample MIS with the proposal changes the noise, not the expectation:
    guided and plain images agree in mean to within their Monte-Carlo error."""
    print_case(f"guided render unbiased (trained={trained})")
    s, cam = _cloud()
    gs = ctx.upload(s)
    g = Guide(ctx, log2_table=16, batch=16384, seed=4, guided_vertices=3, guide_prob=0.5)
    W = H = 32
    if trained:
        f = ctx.film(W, H)
        ctx.render(gs, cam, f, spp=256, max_bounces=64, seed=99, guide=g, train=True)
        assert g.stats()["steps"] >= 1
    spp = 256
    fp, fg = ctx.film(W, H), ctx.film(W, H)
    ctx.render(gs, cam, fp, spp=spp, max_bounces=64, seed=1)
    ctx.render(gs, cam, fg, spp=spp, max_bounces=64, seed=2, guide=g)
    a = fp.read().mean(2) / spp
    b = fg.read().mean(2) / spp
    dlt = (b - a).reshape(-1)
    z = dlt.mean() / (dlt.std() / np.sqrt(dlt.size))
    assert abs(z) < 5.0, (z, a.mean(), b.mean())
    assert abs(b.mean() - a.mean()) < 0.03 * a.mean()
    g.close()


def test_guided_render_trains_online(ctx):
    print_case("guided render trains online")
    s, cam = _cloud()
    gs = ctx.upload(s)
    g = Guide(ctx, log2_table=16, batch=16384, seed=5)
    f = ctx.film(64, 64)
    ctx.render(gs, cam, f, spp=64, max_bounces=64, seed=3, guide=g, train=True)
    st = g.stats()
    assert st["steps"] >= 2 and st["samples"] >= 2 * 8192, st
    assert np.isfinite(st["last_loss"])
    q = np.array([[0.5, 0.5, 0.5, 0.0, 0.0, 1.0]], np.float32)
    mix = g.eval(q).reshape(8, 5)
    assert np.all(np.isfinite(mix)) and abs(mix[:, 0].sum() - 1) < 1e-5
    g.close()


def test_trained_guide_gains_over_phase_sampling(ctx):
    """What the trained network buys: on the cumulus cloud (160 x 90, the
    cloud workload's camera and 128 bounces), a guide trained online over
    2^19 samples per pixel, then frozen, renders 256-spp images whose relMSE
    against a 32768-spp plain render is lower than that of plain HG sampling
    at the same spp (8 seeds each).  The gain is modest (about 5%): the
    g = 0.877 lobe is already close to the product it learns."""
    print_case("trained guide vs phase sampling at equal spp")
    from narvalengine_amd import configs
    W, H = 160, 90
    wl = configs.cloud512(width=W, height=H)
    gs = ctx.upload(wl.scene)

    def render(spp, seed, **ex):
        f = ctx.film(W, H)
        ctx.render(gs, wl.camera, f, spp=spp, seed=seed, max_bounces=wl.max_bounces, **ex)
        img = f.read() / spp
        f.close()
        return img

    def rel_mse(x, ref):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))

    ref = render(32768, 1000)
    g = Guide(ctx, seed=11, guided_vertices=1, guide_prob=0.25)
    render(1 << 19, 5, guide=g, train=True)
    assert g.stats()["steps"] >= 200
    plain = [rel_mse(render(256, 10 + k), ref) for k in range(8)]
    guided = [rel_mse(render(256, 20 + k, guide=g), ref) for k in range(8)]
    ratio = np.mean(guided) / np.mean(plain)
    print(f"relMSE plain {np.mean(plain):.4e} guided {np.mean(guided):.4e} ratio {ratio:.3f}")
    assert ratio < 0.99, (plain, guided)
    g.close()
    gs.close()


def _oct_centres(nd):
    """Bin-centre directions of the octahedral map that include/ne_gpu.h
    documents for ne_guide_get_cache, in FP32 with one rounding per operation
    (the device's order), so the network sees bit-identical inputs."""
    f = np.float32
    b = np.arange(nd * nd)
    inv = f(1) / f(nd)
    u = ((b % nd).astype(f) + f(0.5)) * inv
    v = ((b // nd).astype(f) + f(0.5)) * inv
    px, py = f(2) * u - f(1), f(2) * v - f(1)
    pz = (f(1) - np.abs(px)) - np.abs(py)
    fold = pz < 0
    ox = (f(1) - np.abs(py)) * np.where(px >= 0, f(1), f(-1))
    oy = (f(1) - np.abs(px)) * np.where(py >= 0, f(1), f(-1))
    px, py = np.where(fold, ox, px).astype(f), np.where(fold, oy, py).astype(f)
    n = np.sqrt((px * px + py * py) + pz * pz).astype(f)
    return np.stack([px / n, py / n, pz / n], 1).astype(f)


def test_tabulated_proposal_is_the_network_at_entry_centres(ctx):
    """Every tabulated entry equals the network's mixture at its cell centre and
    direction-bin centre (the same encoder and MFMA tiles evaluate both)."""
    print_case("guide tabulation vs network")
    nc, nd = 5, 3
    g = Guide(ctx, log2_table=LOG2T, batch=4096, seed=1, cache_res=nc, cache_dirs=nd)
    table, mlp = _random_params(g, 8)
    g.set_params(table, mlp)
    tab = g.cache()
    assert tab.shape == (nc ** 3 * nd * nd, 40)
    e = np.arange(len(tab))
    cell, b = e // (nd * nd), e % (nd * nd)
    f = np.float32
    inv = f(1) / f(nc)
    p = np.stack([((cell % nc).astype(f) + f(0.5)) * inv, ((cell // nc % nc).astype(f) + f(0.5)) * inv,
                  ((cell // (nc * nc)).astype(f) + f(0.5)) * inv], 1)
    d = _oct_centres(nd)[b]
    net = g.eval(np.concatenate([p, d], 1).astype(np.float32))
    np.testing.assert_array_equal(tab, net)
    # a parameter change re-tabulates on the next read
    g.set_params(table * 0.5, mlp)
    assert np.abs(g.cache() - tab).max() > 1e-4
    g.close()


def _trained_cached_guide(ctx, gs, cam, **kw):
    g = Guide(ctx, log2_table=16, batch=16384, seed=4, cache_res=8, cache_dirs=4, **kw)
    f = ctx.film(32, 32)
    ctx.render(gs, cam, f, spp=128, max_bounces=64, seed=99, guide=g, train=True)
    assert g.stats()["steps"] >= 1
    f.close()
    return g


def test_guided_megakernel_traces_the_guided_wavefront_paths(ctx):
    """With the tabulated proposal the megakernel and the wavefront make the same
    draws and the same arithmetic at guided vertices: same image."""
    print_case("guided megakernel vs guided wavefront")
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT
    s, cam = _cloud()
    gs = ctx.upload(s)
    g = _trained_cached_guide(ctx, gs, cam, guided_vertices=4, guide_prob=0.5)
    out = []
    for mode in (NE_MODE_WAVEFRONT, NE_MODE_MEGAKERNEL):
        f = ctx.film(40, 24)
        ctx.render(gs, cam, f, spp=16, max_bounces=64, seed=13, guide=g, mode=mode)
        out.append(f.read())
        f.close()
    assert out[0].any()
    np.testing.assert_allclose(out[1], out[0], rtol=1e-6, atol=1e-9)
    g.close()


def test_guided_batch_job_equals_its_single_job_render(ctx):
    """Guided + batched: a guided job in a batch with plain jobs is bitwise the
    guided job batched alone, and equals the guided megakernel render of it
    up to the film's fixed-point quantisation."""
    print_case("guided batch job vs single-job render")
    from narvalengine_amd import Batch
    from narvalengine_amd.abi import NE_MODE_MEGAKERNEL
    s, cam = _cloud()
    gs = ctx.upload(s)
    g = _trained_cached_guide(ctx, gs, cam, guided_vertices=4, guide_prob=0.5)
    import scenes
    other = ctx.upload(scenes.small_volume(res=16, maj_res=2))
    jobs = [dict(scene=other, camera=scenes.VOLUME_CAMERA, width=24, height=16, spp=8, max_bounces=16, seed=3, job=1),
            dict(scene=gs, camera=cam, width=40, height=24, spp=12, max_bounces=64, seed=5, job=2, guide=g),
            dict(scene=other, camera=scenes.VOLUME_CAMERA, width=16, height=16, spp=4, max_bounces=16, seed=4, job=3)]
    b = Batch(ctx, jobs)
    b.run((8, 8, 4))
    got = b.read(1)
    alone = Batch(ctx, [jobs[1]])
    alone.run((16, 4, 2))
    np.testing.assert_array_equal(alone.read(0), got)
    f = ctx.film(40, 24)
    ctx.render(gs, cam, f, spp=12, max_bounces=64, seed=5, job=2, guide=g, mode=NE_MODE_MEGAKERNEL)
    np.testing.assert_allclose(got, f.read(), rtol=1e-6, atol=12 * 2.0 ** -30 * 4)
    # the plain jobs are untouched by the guided one
    plain = Batch(ctx, [jobs[0]])
    plain.run((8, 8, 4))
    np.testing.assert_array_equal(plain.read(0), b.read(0))
    for x in (b, alone, plain):
        x.close()
    g.close()


def test_guided_job_in_the_wavefront_executor(ctx):
    """Guided + batched through the wavefront executor: the guided job equals
    its frozen guided wavefront render up to the fixed-point film, and the
    plain jobs beside it are bitwise those of a batch without it."""
    print_case("guided batch job, wavefront executor")
    from narvalengine_amd import Batch
    from narvalengine_amd.abi import NE_BATCH_WAVEFRONT, NE_MODE_WAVEFRONT
    import scenes
    s, cam = _cloud()
    gs = ctx.upload(s)
    g = _trained_cached_guide(ctx, gs, cam, guided_vertices=4, guide_prob=0.5)
    other = ctx.upload(scenes.small_volume(res=16, maj_res=2))
    jobs = [dict(scene=other, camera=scenes.VOLUME_CAMERA, width=24, height=16, spp=8, max_bounces=16, seed=3, job=1),
            dict(scene=gs, camera=cam, width=40, height=24, spp=12, max_bounces=64, seed=5, job=2, guide=g),
            dict(scene=other, camera=scenes.VOLUME_CAMERA, width=16, height=16, spp=4, max_bounces=16, seed=4, job=3)]
    wf = (8, 8, 4, NE_BATCH_WAVEFRONT)
    b = Batch(ctx, jobs)
    b.run(wf)
    f = ctx.film(40, 24)
    ctx.render(gs, cam, f, spp=12, max_bounces=64, seed=5, job=2, guide=g, mode=NE_MODE_WAVEFRONT)
    np.testing.assert_allclose(b.read(1), f.read(), rtol=1e-6, atol=12 * 2.0 ** -30 * 4)
    plain = Batch(ctx, [jobs[0], jobs[2]])
    plain.run(wf)
    np.testing.assert_array_equal(plain.read(0), b.read(0))
    np.testing.assert_array_equal(plain.read(1), b.read(2))
    for x in (b, plain):
        x.close()
    g.close()


def test_tabulated_guided_render_is_unbiased(ctx):
    """The tabulated proposal leaves the expectation unchanged, as the
    per-vertex network does (the first four vertices of every path guided)."""
    print_case("tabulated guided render unbiased")
    s, cam = _cloud()
    gs = ctx.upload(s)
    g = _trained_cached_guide(ctx, gs, cam, guided_vertices=4, guide_prob=0.5)
    W = H = 32
    spp = 256
    fp, fg = ctx.film(W, H), ctx.film(W, H)
    ctx.render(gs, cam, fp, spp=spp, max_bounces=64, seed=1)
    ctx.render(gs, cam, fg, spp=spp, max_bounces=64, seed=2, guide=g)
    a = fp.read().mean(2) / spp
    b = fg.read().mean(2) / spp
    dlt = (b - a).reshape(-1)
    z = dlt.mean() / (dlt.std() / np.sqrt(dlt.size))
    assert abs(z) < 5.0, (z, a.mean(), b.mean())
    assert abs(b.mean() - a.mean()) < 0.03 * a.mean()
    g.close()
