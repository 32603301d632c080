"""Pins of the guide's CPU restatement (oracle/guide_ref.py): the vMF density
against its closed forms, the hash encoding against exact cases, and the
analytic loss gradient against central finite differences in FP64."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import guide_ref as G  # noqa: E402

LOG2T = 10
RES = [2, 3, 4, 5, 6, 8, 9, 11, 13, 16, 19, 23, 27, 32, 38, 45]


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    table = rng.normal(0, 0.5, G.LEVELS * (1 << LOG2T) * G.FEAT)
    mlp = np.zeros(G.N_PARAMS)
    W, b = G.unpack(mlp)
    for l in range(G.LAYERS):
        used_in = 35 if l == 0 else 64
        used_out = 40 if l == G.LAYERS - 1 else 64
        mlp[G.W_OFF[l]:G.W_OFF[l + 1]] = 0
        w = np.zeros((G.IN[l], G.OUT[l]))
        w[:used_in, :used_out] = rng.normal(0, scale, (used_in, used_out)) * np.sqrt(2 / used_in) / scale
        mlp[G.W_OFF[l]:G.W_OFF[l + 1]] = w.reshape(-1)
        bb = np.zeros(G.OUT[l])
        bb[:used_out] = rng.normal(0, 0.1, used_out)
        mlp[G.B_OFF[l]:G.B_OFF[l + 1]] = bb
    return table, mlp


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    r = rng.uniform(0.1, 2.0, n)
    return np.concatenate([p, d, w, r[:, None]], 1)


@pytest.mark.parametrize("kappa", [1e-3, 0.1, 1.0, 10.0, 300.0, 2980.0])
def test_vmf_density_is_normalised_and_matches_the_sinh_form(kappa):
    # integral over the sphere: 2 pi int_{-1}^{1} p(c) dc, nodes packed near c = 1
    t = np.concatenate([[0.0], np.geomspace(1e-12, 2.0, 200000)])
    c = 1.0 - t
    p = G.vmf_pdf(kappa, c)
    assert abs(2 * np.pi * np.trapezoid(p[::-1], c[::-1]) - 1) < 1e-5
    if kappa < 300:  # closed form kappa / (4 pi sinh kappa) exp(kappa c)
        np.testing.assert_allclose(p[:1000], kappa / (4 * np.pi * np.sinh(kappa)) * np.exp(kappa * c[:1000]),
                                   rtol=1e-10)


def test_hash_encoding_reads_table_entries_at_grid_vertices():
    table, _ = _params(1)
    T = 1 << LOG2T
    tab = table.reshape(G.LEVELS, T, G.FEAT)
    # level 0 (res 2, dense): vertex (1, 2, 0) of the 3^3 grid
    p = np.array([[0.5, 1.0, 0.0]])
    e = G.encode(p, table, RES, LOG2T)
    idx = 1 + 3 * (2 + 3 * 0)
    np.testing.assert_allclose(e[0, 0:2], tab[0, idx], rtol=1e-12)
    # a hashed level (res 45: 46^3 > 1024): vertex (9, 0, 44)
    l = 15
    p = np.array([[9 / 45, 0.0, 44 / 45]])
    e = G.encode(p, table, RES, LOG2T)
    h = (9 ^ ((0 * 2654435761) & 0xFFFFFFFF) ^ ((44 * 805459861) & 0xFFFFFFFF)) & (T - 1)
    np.testing.assert_allclose(e[0, 2 * l:2 * l + 2], tab[l, h], rtol=1e-6)
    # trilinear: the encoding is affine along an axis inside a cell of a dense level
    xs = np.linspace(0.51, 0.99, 7)[:, None] * np.array([[1, 0, 0]]) / 2 + np.array([[0.0, 0.25, 0.25]])
    e = G.encode(xs, table, RES, LOG2T)[:, 0]
    assert np.abs(np.diff(e, 2)).max() < 1e-6  # FP32 rounding of p * res, as on the device


def test_analytic_gradient_matches_finite_differences():
    table, mlp = _params(2)
    s = _samples(24, 3)
    loss, gt, gm = G.loss_and_grad(s, table, mlp, RES, LOG2T, quantize=False)
    rng = np.random.default_rng(4)
    h = 1e-6

    def L(t, m):
        return G.loss_and_grad(s, t, m, RES, LOG2T, quantize=False)[0]

    # MLP parameters: weights of every layer and biases (used entries only)
    picks = []
    for l in range(G.LAYERS):
        used_in = 35 if l == 0 else 64
        used_out = 40 if l == G.LAYERS - 1 else 64
        for _ in range(4):
            i, o = rng.integers(used_in), rng.integers(used_out)
            picks.append(G.W_OFF[l] + i * G.OUT[l] + o)
        picks.append(G.B_OFF[l] + rng.integers(used_out))
    for k in picks:
        mp, mm = mlp.copy(), mlp.copy()
        mp[k] += h
        mm[k] -= h
        fd = (L(table, mp) - L(table, mm)) / (2 * h)
        assert abs(fd - gm[k]) < 1e-6 + 1e-5 * abs(fd), (k, fd, gm[k])
    # hash features that the batch touches
    touched = np.nonzero(gt)[0]
    for k in rng.choice(touched, 10, replace=False):
        tp, tm = table.copy(), table.copy()
        tp[k] += h
        tm[k] -= h
        fd = (L(tp, mlp) - L(tm, mlp)) / (2 * h)
        assert abs(fd - gt[k]) < 1e-7 + 1e-5 * abs(fd), (k, fd, gt[k])
    assert np.isfinite(loss)


def test_vmf_sampler_inverse_cdf_matches_the_density():
    """cos theta = 1 + log1p((1 - u) expm1(-2 kappa)) / kappa is the inverse CDF
    of the vMF density in cos theta, from tiny to large kappa."""
    u = (np.arange(20000) + 0.5) / 20000
    for kappa in (4.5e-5, 1e-2, 1.0, 30.0, 2980.0):
        c = 1 + np.log1p((1 - u) * np.expm1(-2 * kappa)) / kappa
        # CDF of cos theta: 2 pi int_{-1}^{c} p = (exp(kappa (c - 1)) - exp(-2 kappa)) / (1 - exp(-2 kappa))
        cdf = (np.exp(kappa * (c - 1)) - np.exp(-2 * kappa)) / -np.expm1(-2 * kappa)
        np.testing.assert_allclose(cdf, u, atol=2e-9 if kappa > 1e-3 else 2e-6)
