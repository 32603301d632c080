"""Environment map (InfiniteAreaLight, src/lights/InfiniteAreaLight.h:8-130;
Distribution1D/2D, src/utils/Sampling.h:5-112) in the oracle, pinned by
closed forms: the reference's Distribution2D against brute force, the
corrected sampler's histogram against its pdf, the pdf's integral over the
sphere, the lookup frame against InfiniteAreaLight::Le's, and renders against
the constant environment and a quadrature."""
import numpy as np
import pytest

import oracle as O
from narvalengine_amd.abi import NE_ENV_CONSTANT
from narvalengine_amd.scene import Camera, Scene


def _map(H, W, seed=0):
    rng = np.random.default_rng(seed)
    m = rng.gamma(0.6, 1.0, (H, W, 3)).astype(np.float32)
    m[H // 3:H // 3 + 2, W // 4:W // 4 + 3] *= 40.0  # a bright patch
    m[:, W - 2] = 0.0  # black columns
    return m


def _world(m):
    s = Scene()
    s.set_env_map(m)
    return O.World(s), s


def _lum(m):
    return 0.2126 * m[..., 0] + 0.7152 * m[..., 1] + 0.0722 * m[..., 2]


def test_reference_distribution2d_by_brute_force():
    """The reference's Distribution2D on a square function (where its
    conditional size, height, equals the width): samples land in cell (i, j)
    with probability f(i, j) / sum f, and the pdf is f(i, j) / mean f."""
    rng = np.random.default_rng(3)
    H = W = 8
    f = rng.uniform(0, 1, (H, W)).astype(np.float32)
    f[2, 5] = 0.0
    u2 = rng.uniform(0, 1, (200000, 2)).astype(np.float32)
    uv, pdf = O.ref_dist2d(f, H, u2)
    i = np.minimum((uv[:, 0] * W).astype(int), W - 1)
    j = np.minimum((uv[:, 1] * H).astype(int), H - 1)
    counts = np.bincount(j * W + i, minlength=H * W).reshape(H, W)
    p = f / f.sum()
    assert counts[2, 5] == 0
    assert np.all(np.abs(counts / len(u2) - p) < 5 * np.sqrt(p * (1 - p) / len(u2)) + 1e-6)
    np.testing.assert_allclose(pdf, f[j, i] / f.mean(), rtol=2e-5)


def test_corrected_sampler_histogram_matches_its_pdf():
    H, W = 16, 32  # not square: each row's conditional spans its W texels
    m = _map(H, W)
    w, _ = _world(m)
    rng = np.random.default_rng(4)
    n = 400000
    d, pdf, le = w.env_sample(rng.uniform(0, 1, (n, 2)))
    # every sample lies in a texel whose radiance it returns, with that texel's pdf
    p2, le2 = w.env_eval(d)
    ok = np.abs(pdf - p2) <= 1e-4 * p2 + 1e-9
    assert ok.mean() > 0.999
    np.testing.assert_array_equal(le[ok], le2[ok])
    # texel frequencies: func / sum(func), func = luminance * sin(theta of the row centre)
    th = np.arccos(np.clip(d[:, 2], -1, 1))
    ph = np.mod(np.arctan2(-d[:, 0], d[:, 1]), 2 * np.pi)
    i = np.minimum((ph / (2 * np.pi) * W).astype(int), W - 1)
    j = np.minimum((th / np.pi * H).astype(int), H - 1)
    counts = np.bincount(j * W + i, minlength=H * W).reshape(H, W) / n
    func = _lum(m.astype(np.float64)) * np.sin(np.pi * (np.arange(H) + 0.5) / H)[:, None]
    p = func / func.sum()
    assert np.all(np.abs(counts - p) < 5 * np.sqrt(p * (1 - p) / n) + 2e-6)
    assert counts[:, W - 2].sum() == 0  # black texels are never drawn


def test_env_pdf_integrates_to_one_and_le_uses_the_reference_frame():
    H, W = 12, 20
    m = _map(H, W, seed=5)
    w, _ = _world(m)
    # midpoint quadrature over (theta, phi): sum p(d) sin(theta) dtheta dphi = 1
    k = 600
    th = (np.arange(k) + 0.5) * np.pi / k
    ph = (np.arange(2 * k) + 0.5) * np.pi / k
    T, P = np.meshgrid(th, ph, indexing="ij")
    # InfiniteAreaLight::Le frame: local (x, y, z) = (d.y, -d.x, d.z), phi = atan2(y, x), theta = acos(z)
    d = np.stack([-np.sin(T) * np.sin(P), np.sin(T) * np.cos(P), np.cos(T)], -1).reshape(-1, 3)
    pdf, le = w.env_eval(d)
    integral = (pdf.reshape(T.shape) * np.sin(T)).sum() * (np.pi / k) ** 2
    assert abs(integral - 1) < 2e-3, integral
    i = np.minimum((P / (2 * np.pi) * W).astype(int), W - 1).reshape(-1)
    j = np.minimum((T / np.pi * H).astype(int), H - 1).reshape(-1)
    inner = (np.abs(P.reshape(-1) * W / (2 * np.pi) - np.round(P.reshape(-1) * W / (2 * np.pi))) > 1e-3) & \
            (np.abs(T.reshape(-1) * H / np.pi - np.round(T.reshape(-1) * H / np.pi)) > 1e-3)
    np.testing.assert_array_equal(le[inner], m[j[inner], i[inner]])


def test_constant_map_renders_like_the_constant_environment():
    """A white furnace medium under a constant map equals the constant
    environment: importance sampling changes the noise, not the mean."""
    def furnace(env):
        s = Scene()
        s.set_volume(np.full((8, 8, 8), 1.0, np.float32), 3.0, (1.0,) * 3, (0.0,) * 3, 0.6, maj_res=2)
        env(s)
        cam = Camera(position=(0.0, 0.1, -2.0), look_at=(0, 0, 0), up=(0, 1, 0), vfov=40.0, aspect=1.0)
        return O.World(s).render(cam, 8, 8, 128, 128, seed=2) / 128
    a = furnace(lambda s: s.set_env(NE_ENV_CONSTANT, c0=(0.7, 0.7, 0.7)))
    b = furnace(lambda s: s.set_env_map(np.full((8, 16, 3), 0.7, np.float32)))
    assert abs(a.mean() - 0.7) < 0.01 and abs(b.mean() - 0.7) < 0.01, (a.mean(), b.mean())


def test_direct_light_from_a_map_patch_matches_quadrature():
    """A black map with one bright texel block above a glossy floor: the MIS
    estimate at a floor point equals the quadrature of f Le cos over the
    patch's solid angle."""
    H, W = 16, 32
    m = np.zeros((H, W, 3), np.float32)
    m[3:5, 6:9] = (30.0, 20.0, 10.0)
    s = Scene()
    albedo, rough, metal = (0.7, 0.6, 0.5), 0.5, 0.0
    floor = s.add_glossy(albedo, rough, metal)
    s.add_rectangle(floor, position=(0, 0, 0), rotation=(90, 0, 0), scale=(50, 50, 1))
    s.set_env_map(m)
    cam = Camera(position=(0.0, 1.0, -1.0), look_at=(0.0, 0.0, 0.0), up=(0, 1, 0), vfov=0.005, aspect=1.0)
    w = O.World(s, fp64=True)
    vals = [w.render(cam, 1, 1, 20000, 1, seed=1, sample_offset=b * 20000)[0, 0] / 20000 for b in range(8)]
    mean, se = np.mean(vals, 0), np.std(vals, 0, ddof=1) / np.sqrt(8)
    # quadrature over the patch in (theta, phi) of the map frame
    k = 200
    th = (3 + 2 * (np.arange(k) + 0.5) / k) * np.pi / H
    ph = (6 + 3 * (np.arange(k) + 0.5) / k) * 2 * np.pi / W
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([-np.sin(T) * np.sin(P), np.sin(T) * np.cos(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (2 * np.pi / H / k) * (3 * 2 * np.pi / W / k)).reshape(-1)
    up = d[:, 1] > 0
    wo = np.array([0.0, 1.0, -1.0]) / np.sqrt(2)
    to_local = lambda v: np.stack([v[..., 0], v[..., 2], v[..., 1]], -1)
    n = int(up.sum())
    f = np.zeros((n, 3), np.float32)
    pdf, smp = np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
    mat = np.array([*albedo, rough, metal], np.float32)
    O.lib().neo_glossy_c(n, O.fp(np.tile(mat, (n, 1))), O.fp(np.tile(to_local(wo), (n, 1)).astype(np.float32)),
                         O.fp(to_local(d[up]).astype(np.float32)), O.fp(np.zeros((n, 3), np.float32)), O.fp(f),
                         O.fp(pdf), O.fp(smp))
    Le = np.array([30.0, 20.0, 10.0])
    Ld = (f * (d[up, 1] * dw[up])[:, None]).sum(0) * Le
    assert Ld.min() > 0
    assert (np.abs(mean - Ld) < 5 * se + 3e-3 * Ld).all(), (mean, Ld, se)
