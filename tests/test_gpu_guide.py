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
