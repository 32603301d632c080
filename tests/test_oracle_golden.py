"""Pins the oracle against the reference's own golden vectors and fixtures
(unitTests/tests.cpp, resources/tests/avgHGCos.txt) and the published
Philox-4x32-10 known-answer vectors.  CPU only."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
EPS3 = 1e-3  # floatEqualsEpsilon tolerance, unitTests/tests.cpp:27-43


# ---- Philox-4x32-10 known-answer tests (Random123 kat_vectors) -------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, expect):
    assert list(O.philox(ctr, key)) == expect


def test_rng_stream_layout():
    # uniform i of a path = top 24 bits of Philox word i % 4 of block i // 4
    u = O.rng_uniforms(seed=0x1234_5678_9abc_def0, pixel=7, sample=3, job=2, n=8)
    key = [0x9abcdef0, 0x12345678]
    w0 = O.philox([7, 3, 0, 2], key)
    w1 = O.philox([7, 3, 1, 2], key)
    words = np.concatenate([w0, w1])
    np.testing.assert_array_equal(u, (words >> 8).astype(np.float32) * np.float32(2.0 ** -24))
    assert (u >= 0).all() and (u < 1).all()  # Math.random, tests.cpp:48-57


# ---- Triangle.intersection, unitTests/tests.cpp:300-347 --------------------
def test_triangle_intersection_goldens():
    tri = [0, 0, 0, 0, 1, 0, 1, 0, 0]
    rays = [([0, 0, 0], [0, 0, 1], True, 0.0, [0, 0, 0]),
            ([0, 0, -1], [0, 0, 1], True, 1.0, [0, 0, 0]),
            ([0, -1, 0], [0, 1, 0], False, None, None),
            ([0.25, 0.25, -1], [0, 0, 1], True, 1.0, [0.25, 0.25, 0]),
            ([0, 0, 1], [0, 0, 1], False, None, None)]
    for o, d, hit, t, p in rays:
        h, tt, _ = O.triangle([*o, *d], tri)
        assert bool(h[0]) == hit
        if hit:
            assert tt[0] == t  # exact equality, as the reference test
            np.testing.assert_allclose(np.array(o) + tt[0] * np.array(d), p, atol=EPS3)


# ---- AABB.intersection, unitTests/tests.cpp:352-404 ------------------------
def test_aabb_intersection_goldens():
    box = [-0.5, -0.5, -0.5, 0.5, 0.5, 0.5]
    cases = [([0, 0, -1.0], True, 0.5, 1.5, [0, 0, -0.5]),
             ([0, 0, -0.5], True, 0.0, 1.0, [0, 0, -0.5]),
             ([0, -0.5, -0.5], True, 0.0, 1.0, [0, -0.5, -0.5]),
             ([0, 0.0, 0.0], True, -0.5, 0.5, [0, 0, -0.5])]
    for o, hit, tn, tf, p in cases:
        h, a, b, pt = O.aabb([*o, 0, 0, 1], box)
        assert h[0] == 1
        assert abs(a[0] - tn) < EPS3 and abs(b[0] - tf) < EPS3
        np.testing.assert_allclose(pt[0], p, atol=EPS3)
    h, a, b, _ = O.aabb([0, 0.55, 0.0, 0, 0, 1], box)
    assert h[0] == 0 and a[0] == np.inf and b[0] == -np.inf


# ---- Math.generateOrthonormalCS / conversionBetweenLCSandWCS ---------------
def test_orthonormal_frame_goldens():
    n = O.f([[0, 1, 0], [0, -1, 0]])
    v = np.zeros((2, 3), np.float32)
    u = np.zeros((2, 3), np.float32)
    O.lib().neo_onb(2, O.fp(n), O.fp(v), O.fp(u))
    np.testing.assert_allclose(u[0], [-1, 0, 0], atol=EPS3)
    np.testing.assert_allclose(v[0], [0, 0, -1], atol=EPS3)
    np.testing.assert_allclose(u[1], [-1, 0, 0], atol=EPS3)
    np.testing.assert_allclose(v[1], [0, 0, 1], atol=EPS3)


def test_lcs_wcs_goldens():
    ns, ss, ts = [[0, 1, 0]], [[1, 0, 0]], [[0, 0, 1]]
    lcs = O.vec_op("neo_to_lcs", 1, [[0, 1, 0]], ns, ss, ts, out_shape=(1, 3))
    wcs = O.vec_op("neo_to_world", 1, [[0, 0, 1]], ns, ss, ts, out_shape=(1, 3))
    np.testing.assert_allclose(lcs[0], [0, 0, 1], atol=EPS3)
    np.testing.assert_allclose(wcs[0], [0, 1, 0], atol=EPS3)


# ---- Math.sampleUnitSphere, unitTests/tests.cpp:103-136 --------------------
def test_sample_unit_sphere_goldens():
    cases = []
    for th in (0, 90, 180, 270, 360):
        cases.append(((th / 360, 0.0), [0, 0, 1]))
        cases.append(((th / 360, 1.0), [0, 0, -1]))
    eq = {0: [1, 0, 0], 90: [0, 1, 0], 180: [-1, 0, 0], 270: [0, -1, 0], 360: [1, 0, 0]}
    for th, p in eq.items():
        cases.append(((th / 360, 0.5), p))
    e = np.array([c[0] for c in cases], np.float32)
    out = O.vec_op("neo_sample_unit_sphere", len(e), e, out_shape=(len(e), 3))
    np.testing.assert_allclose(out, [c[1] for c in cases], atol=EPS3)


# ---- Triangle.area / contains / barycentricCoordinates (tests.cpp:650-864) --
def test_triangle_area_goldens():
    tris = [[0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 0, 0, 2, 0, 2, 0, 0], [0, 0, 0, 0, 2, 0, 0, 0, 2],
            [0, 0, 0, 0, 2, 0, -2, 0, 0], [0, 0, 0, 0, 2, 0, 0, 0, -2], [15, 0, 15, 15, 0, 16, 15, 1, 15],
            [-15, 0, -15, -15, 0, -16, -15, -1, -15]]
    a = O.vec_op("neo_triangle_area", len(tris), tris, out_shape=(len(tris),))
    np.testing.assert_allclose(a, [0.5, 2, 2, 2, 2, 0.5, 0.5], atol=EPS3)


def test_triangle_contains_goldens():
    t1 = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    t2 = [0, 0, 0, 1000, 0, 0, 0, 1000, 0]
    cases = [(t1, [0, 0, 0], 1), (t1, [1, 0, 0], 1), (t1, [0, 1, 0], 1), (t1, [0.3, 0.3, 0], 1),
             (t1, [0.6, 0.6, 0], 0), (t1, [0.3, 0.3, 0.1], 0), (t1, [-0.3, -0.3, 0.1], 0), (t1, [-0.3, -0.3, 0], 0),
             (t2, [0.3, 0.3, 0.1], 0), (t2, [0.3, 0.3, 0.01], 0), (t2, [0.3, 0.3, 0.001], 0),
             (t2, [0.3, 0.3, 0.0001], 0)]
    p = O.f([c[1] for c in cases])
    t = O.f([c[0] for c in cases])
    out = np.zeros(len(cases), np.int32)
    O.lib().neo_triangle_contains(len(cases), O.fp(p), O.fp(t), O.ip(out))
    assert list(out) == [c[2] for c in cases]


def test_barycentric_goldens():
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    pts = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]
    out = O.vec_op("neo_barycentric", 4, pts, [tri] * 4, out_shape=(4, 3))
    np.testing.assert_allclose(out, [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 1, 1]], atol=EPS3)


# ---- BxDF.HGPhaseFunctionEvalG, unitTests/tests.cpp:141-166 ----------------
def test_hg_eval_mean_cosine_is_g():
    n = 100000
    u = O.rng_uniforms(11, 0, 0, 0, 2 * n).reshape(n, 2)
    s = O.vec_op("neo_sample_unit_sphere", n, u, out_shape=(n, 3))
    inc = np.tile([0, 0, 1.0], (n, 1))
    for g in np.arange(-0.9, 0.9, 0.3, dtype=np.float32):
        out = np.zeros(n, np.float32)
        O.lib().neo_hg_eval(n, float(g), O.fp(O.f(inc)), O.fp(O.f(s)), O.fp(out))
        result = np.mean(out.astype(np.float64) * s[:, 2]) / (1 / (4 * np.pi))
        assert abs(result - g) < 0.05


# ---- BxDF.HGPhaseFunction, unitTests/tests.cpp:171-221 ---------------------
def test_hg_eval_bucket_average():
    n_th, n_ph = 100, 50
    th_step, ph_step = np.float32(2 * np.pi / n_th), np.float32(np.pi / n_ph)
    inc = []
    th = np.float32(0)
    ph = np.float32(0)
    for i in range(n_th):
        for j in range(n_ph):
            # sampleUnitSphere(cos, sin, phi) = (sin(phi) cos, sin(phi) sin, cos(phi)), Math.h:477-479
            inc.append([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)])
            ph = np.float32(ph + ph_step)
        th = np.float32(th + th_step)
    inc = O.f(inc)
    sc = O.f(np.tile([1, 0, 0], (len(inc), 1)))
    out = np.zeros(len(inc), np.float32)
    O.lib().neo_hg_eval(len(inc), 0.1, O.fp(inc), O.fp(sc), O.fp(out))
    assert abs(out.astype(np.float64).mean() - 1 / (4 * np.pi)) < EPS3


# ---- resources/tests/avgHGCos.txt: mean polar angle of HG::sample ----------
def _hg_mean_theta_closed_form(g):
    # samples of the reference sampler follow the HG density mirrored to -z:
    # p(mu) = (1 - g^2) / (2 (1 + g^2 + 2 g mu)^1.5), mu = cos(theta)
    if abs(g) > 0.999:
        return np.pi if g > 0 else 0.0
    mu = np.linspace(-1, 1, 2_000_001)
    p = (1 - g * g) / (2 * (1 + g * g + 2 * g * mu) ** 1.5)
    return float(np.trapezoid(np.arccos(mu) * p, mu) / np.trapezoid(p, mu))


def test_hg_sampler_matches_avgHGCos_fixture():
    fixture = np.loadtxt(GOLDEN / "avgHGCos.txt")
    gs = np.round(np.arange(-1.0, 1.0001, 0.1), 1)
    assert len(fixture) == len(gs) == 21
    n = 200000
    u = O.rng_uniforms(5, 0, 0, 0, 2 * n).reshape(n, 2)
    for g, fx in zip(gs, fixture):
        out = np.zeros((n, 3), np.float32)
        O.lib().neo_hg_sample_local(n, float(g), O.fp(O.f(u)), O.fp(out))
        theta = np.arccos(np.clip(out[:, 2].astype(np.float64), -1, 1))
        cf = _hg_mean_theta_closed_form(g)
        sd = theta.std() + 1e-3
        # oracle sampler vs closed form: 200k samples
        assert abs(theta.mean() - cf) < 5 * sd / np.sqrt(n) + 2e-3, (g, theta.mean(), cf)
        # reference run (1000 samples per g, TestPlayground.h:852-892) vs closed form
        assert abs(fx - cf) < 4 * sd / np.sqrt(1000) + 2e-3, (g, fx, cf)
