"""Homogeneous media (HomogeneousMedia, src/materials/HomogeneousMedia.cpp:5-51)
on the CPU: the oracle's restatement of the reference functions, and the
corrected estimator the device runs, each pinned by closed forms or by an
independent routine (the reference's own ratio tracking on a constant grid)."""
import numpy as np
import pytest

import oracle as O
import scenes
from narvalengine_amd.abi import NE_ENV_CONSTANT
from narvalengine_amd.scene import Camera, Scene, transform


def test_reference_tr_equals_ratio_tracking_on_a_constant_grid():
    """HomogeneousMedia::Tr against GridMedia::Tr (ratio tracking) through a
    grid of ones: two unrelated routines with the same expectation."""
    res, mult, ext = 16, 2.5, 0.8
    grid = np.ones(res ** 3, np.float32)
    n = 100000
    L = 1.0 - 1.0 / res  # stays where the trilinear field of ones is exactly 1
    rays = np.tile([-0.5, 0.1, -0.2, 1.0, 0.0, 0.0], (n, 1)).astype(np.float32)
    tint = np.tile([0.0, L], (n, 1)).astype(np.float32)
    tr = np.zeros(n, np.float32)
    O.lib().neo_ratio_track(res, res, res, O.fp(grid), 1.0, ext * mult, n, O.fp(rays), O.fp(tint), 5, O.fp(tr))
    closed = O.homogeneous_tr([ext] * 3, [mult], [L])[0]
    assert abs(tr.mean() - closed[0]) < 5 * tr.std() / np.sqrt(n) + 1e-4, (tr.mean(), closed)
    # per channel, multiplicative over distance, 1 at zero distance
    e = np.array([[0.3, 1.0, 2.0]] * 3, np.float32)
    a = O.homogeneous_tr(e, [1.5] * 3, [0.2, 0.5, 0.7])
    np.testing.assert_allclose(a[0] * a[1], a[2], rtol=1e-6)
    np.testing.assert_allclose(O.homogeneous_tr(e[:1], [1.5], [0.0]), 1.0)
    assert a[0, 0] > a[0, 1] > a[0, 2]


@pytest.mark.parametrize("dist", [0.4, 1.0, 1.7])
def test_reference_sample_keeps_its_chord_length_division(dist):
    """The reference divides the free-flight distance by the chord length
    (HomogeneousMedia.cpp:26-28) and leaves the density out: a collision
    happens with probability 1 - exp(-avg(ext) dist^2), and the weights are
    scattering / ext on a collision and 1 on escape (grey extinction) -- or 0
    on escape when Tr of the unclamped sampled distance underflows."""
    ss, sa = 0.7, 0.3
    n = 100000
    tint = np.tile([0.5, 0.5 + dist], (n, 1)).astype(np.float32)
    sampled, t, w = O.homogeneous_sample([ss] * 3, [sa] * 3, 4.0, tint, seed=2)
    p = 1 - np.exp(-(ss + sa) * dist * dist)
    assert abs(sampled.mean() - p) < 5 * np.sqrt(p * (1 - p) / n)
    hit = sampled.astype(bool)
    np.testing.assert_allclose(w[hit], ss / (ss + sa), rtol=1e-5)
    esc = w[~hit]
    assert np.all((np.abs(esc - 1.0) < 1e-6) | (esc == 0.0)) and (esc == 0.0).mean() < 1e-3
    assert (t[hit] < dist).all() and np.allclose(t[~hit], dist)


def _chords(scene, rays):
    """Chord length of each ray through the transformed unit cube (FP64 slab test)."""
    inv = scene.volume["inv"]
    o = rays[:, :3].astype(np.float64) @ inv[:3, :3].T + inv[:3, 3]
    d = rays[:, 3:].astype(np.float64) @ inv[:3, :3].T
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (-0.5 - o) / d
        t1 = (0.5 - o) / d
    tn = np.nanmax(np.minimum(t0, t1), axis=1)
    tf = np.nanmin(np.maximum(t0, t1), axis=1)
    tn = np.maximum(tn, 0.0)
    return np.where(tf > tn, tf - tn, 0.0)


def test_corrected_transmittance_and_free_flight_through_a_rotated_box():
    s = Scene()
    ext, mult = 1.0, 2.0
    s.set_homogeneous_volume(mult, (0.6,) * 3, (ext - 0.6,) * 3, 0.3, position=(0.1, 0.2, -0.1),
                             rotation=(20, 35, 10), scale=(0.8, 0.5, 0.6))
    w = O.World(s)
    n = 50000
    rays = scenes.rays_towards(n, target_box=0.3, origin_radius=2.0, seed=7)
    rays[:, :3] += np.array([0.1, 0.2, -0.1], np.float32)
    L = _chords(s, rays)
    assert (L > 0).mean() > 0.9
    tr = w.transmittance(rays, np.full(n, np.inf, np.float32))
    np.testing.assert_allclose(tr, np.exp(-ext * mult * L), rtol=2e-5, atol=1e-7)
    sc, t = w.delta_track(rays, seed=9)
    p = 1 - np.exp(-ext * mult * L)
    assert abs(sc.mean() - p.mean()) < 5 * np.sqrt((p * (1 - p)).sum()) / n
    # a collision lies inside the chord
    assert (sc.astype(bool) <= (L > 0)).all()


def test_white_furnace_in_a_homogeneous_medium():
    """Albedo 1 under a uniform environment of radiance 1, no surfaces: every
    path returns exactly 1 in expectation."""
    s = Scene()
    s.set_homogeneous_volume(3.0, (1.0,) * 3, (0.0,) * 3, 0.7, rotation=(10, 20, 30), scale=(1.2, 0.8, 1.0))
    s.set_env(NE_ENV_CONSTANT, c0=(1.0, 1.0, 1.0))
    cam = Camera(position=(0.0, 0.1, -2.5), look_at=(0, 0, 0), up=(0, 1, 0), vfov=30.0, aspect=1.0)
    w = O.World(s)
    spp = 256
    img = w.render(cam, 8, 8, spp, 256, seed=4) / spp
    assert abs(img.mean() - 1.0) < 0.01, img.mean()


def test_homogeneous_scene_desc_carries_no_grid():
    s = scenes.homogeneous_box()
    d, keep = s.desc()
    assert d.has_volume == 1 and d.vol_homogeneous == 1 and not d.grid
    assert "grid" not in keep
    np.testing.assert_allclose(np.array(d.vol_xf).reshape(4, 4) @ np.array(d.vol_inv).reshape(4, 4), np.eye(4),
                               atol=1e-6)
    M = transform((0, 0.1, 0), (20, 35, 0), (0.8, 0.5, 0.6))
    np.testing.assert_allclose(np.array(d.vol_xf).reshape(4, 4), M, rtol=1e-6, atol=1e-7)
