"""Pins of the oracle's reference-semantics sphere, rectangle and glossy
functions that the reference's own unit tests leave open.

* ``ref_sphere_intersect`` (src/primitives/Sphere.cpp:21-44): distances along
  a ray through the centre, the roots lying on the sphere, the reference's
  acceptance of negative roots, exact tangency, and hit/miss against the
  distance from the centre to the ray's line (brute force in FP64).
* ``ref_rect_intersect`` (src/primitives/Rectangle.cpp:16-76): a hit at a
  known distance, the inclusive bounds and a miss just outside each of them,
  the flat axis left unchecked (``getSize`` zeroes it), near-parallel and
  backward rays, and hit/miss against an FP64 plane intersection.
* ``ref_glossy_eval`` (src/core/GlossyBSDF.cpp:26-48): the value at normal
  incidence (D = 1/(pi a^2), G = 1, F = 0.04), reciprocity of f / cos(theta_o),
  the diffuse share (1 - F)(1 - metallic) albedo / pi, and the specular lobe's
  directional albedo against an independent half-vector quadrature.

CPU only."""
import numpy as np
import pytest

import oracle as O


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


# ---- Sphere::intersect --------------------------------------------------------
def test_sphere_ray_through_the_centre_from_outside():
    rng = np.random.default_rng(1)
    n = 2000
    c = rng.uniform(-1, 1, (n, 3))
    r = rng.uniform(0.05, 1.0, n)
    u = _unit(rng.normal(size=(n, 3)))
    D = r + rng.uniform(0.2, 4.0, n)
    s = rng.uniform(0.3, 3.0, n)  # direction length: Sphere::intersect does not normalise it
    rays = np.concatenate([c + D[:, None] * u, -u * s[:, None]], axis=1)
    hit, tn, tf = O.sphere(rays, np.concatenate([c, r[:, None]], axis=1))
    assert hit.all()
    np.testing.assert_allclose(tn, (D - r) / s, rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(tf, (D + r) / s, rtol=2e-5, atol=2e-6)


def test_sphere_from_inside_keeps_the_negative_root():
    """The reference returns true with tNear < 0 when the origin is inside."""
    rng = np.random.default_rng(2)
    n = 2000
    c = rng.uniform(-1, 1, (n, 3))
    r = rng.uniform(0.2, 1.0, n)
    o = c + _unit(rng.normal(size=(n, 3))) * (r * rng.uniform(0, 0.95, n))[:, None]
    d = _unit(rng.normal(size=(n, 3)))
    hit, tn, tf = O.sphere(np.concatenate([o, d], axis=1), np.concatenate([c, r[:, None]], axis=1))
    assert hit.all()
    assert (tn < 0).all() and (tf > 0).all()
    for t in (tn, tf):  # both roots lie on the sphere
        p = o + t.astype(np.float64)[:, None] * d
        np.testing.assert_allclose(np.linalg.norm(p - c, axis=1), r, rtol=1e-4, atol=1e-5)


def test_sphere_behind_the_ray_still_reports_a_hit():
    """Both roots negative: the reference still returns true (tNear = min root)."""
    sph = np.array([[0.0, 0.0, 0.0, 1.0]] * 3, np.float32)
    rays = np.array([[0, 0, 5, 0, 0, 1], [0.5, 0, 3, 0, 0, 2], [0, 0, 5, 0, 0, -1]], np.float32)
    hit, tn, tf = O.sphere(rays, sph)
    assert hit.tolist() == [1, 1, 1]
    assert (tn[0], tf[0]) == (-6.0, -4.0)
    np.testing.assert_allclose([tn[1], tf[1]], [(-3 - np.sqrt(0.75)) / 2, (-3 + np.sqrt(0.75)) / 2], rtol=1e-6)
    assert (tn[2], tf[2]) == (4.0, 6.0)


def test_sphere_tangent_and_near_tangent_rays():
    sph = np.array([[0.0, 0.0, 0.0, 1.0]] * 4, np.float32)
    rays = np.array([[1, 0, -5, 0, 0, 1],          # tangent: discriminant exactly 0
                     [1.5, 0, -5, 0, 0, 1],        # misses by 0.5
                     [1.001, 0, -5, 0, 0, 1],      # just outside
                     [0.999, 0, -5, 0, 0, 1]], np.float32)
    hit, tn, tf = O.sphere(rays, sph)
    assert hit.tolist() == [1, 0, 0, 1]
    assert tn[0] == 5.0 and tf[0] == 5.0


def test_sphere_hits_where_the_line_passes_within_the_radius():
    rng = np.random.default_rng(3)
    n = 20000
    c = rng.uniform(-1, 1, (n, 3))
    r = rng.uniform(0.05, 0.8, n)
    o = rng.uniform(-2, 2, (n, 3))
    d = _unit(c + rng.normal(0, 0.5, (n, 3)) - o)
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    sph = np.concatenate([c, r[:, None]], axis=1).astype(np.float32)
    hit, tn, tf = O.sphere(rays, sph)
    o64, d64, c64, r64 = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64), sph[:, :3].astype(
        np.float64), sph[:, 3].astype(np.float64)
    oc = c64 - o64
    t_closest = np.sum(oc * d64, axis=1) / np.sum(d64 * d64, axis=1)
    dist = np.linalg.norm(o64 + t_closest[:, None] * d64 - c64, axis=1)
    clear = np.abs(dist - r64) > 1e-4 * np.maximum(r64, 1)
    assert clear.mean() > 0.99
    np.testing.assert_array_equal(hit[clear].astype(bool), (dist <= r64)[clear])
    assert 0.2 < hit.mean() < 0.8
    m = hit.astype(bool)
    for t in (tn[m], tf[m]):
        p = o64[m] + t.astype(np.float64)[:, None] * d64[m]
        np.testing.assert_allclose(np.linalg.norm(p - c64[m], axis=1), r64[m], rtol=2e-4, atol=2e-5)


# ---- Rectangle::intersect -------------------------------------------------------
def _rect(vA, vB, nrm):
    return np.array([*vA, *vB, *nrm], np.float32)


UNIT_QUAD = _rect((-0.5, -0.5, 0.0), (0.5, 0.5, 0.0), (0.0, 0.0, -1.0))


def test_rectangle_hit_at_a_known_distance():
    rng = np.random.default_rng(4)
    n = 2000
    p = np.concatenate([rng.uniform(-0.49, 0.49, (n, 2)), np.zeros((n, 1))], axis=1)
    d = _unit(np.concatenate([rng.uniform(-0.5, 0.5, (n, 2)), rng.choice([-1.0, 1.0], (n, 1))], axis=1))
    t_true = rng.uniform(0.1, 5.0, n)
    o = p - t_true[:, None] * d   # from either side: the test is two-sided
    hit, t = O.rectangle(np.concatenate([o, d], axis=1), np.tile(UNIT_QUAD, (n, 1)))
    assert hit.all()
    np.testing.assert_allclose(t, t_true, rtol=1e-5, atol=1e-6)


def test_rectangle_bounds_are_inclusive_and_a_step_outside_misses():
    eps = 1e-3
    pts = [(0.5, 0.0), (-0.5, 0.0), (0.0, 0.5), (0.0, -0.5), (0.5, 0.5),
           (0.5 + eps, 0.0), (-0.5 - eps, 0.0), (0.0, 0.5 + eps), (0.0, -0.5 - eps), (0.2, 0.1)]
    rays = np.array([[x, y, -3.0, 0.0, 0.0, 1.0] for x, y in pts], np.float32)
    hit, t = O.rectangle(rays, np.tile(UNIT_QUAD, (len(pts), 1)))
    assert hit.tolist() == [1, 1, 1, 1, 1, 0, 0, 0, 0, 1]
    assert (t[hit == 1] == 3.0).all()
    # corners need not be ordered: bounds are the componentwise min/max
    q = _rect((2.0, 3.0, 0.25), (1.0, 1.0, 0.25), (0.0, 0.0, -1.0))
    rays = np.array([[1.5, 2.0, -1, 0, 0, 1], [2.5, 2.0, -1, 0, 0, 1], [1.5, 3.5, -1, 0, 0, 1]], np.float32)
    hit, t = O.rectangle(rays, np.tile(q, (3, 1)))
    assert hit.tolist() == [1, 0, 0] and t[0] == 1.25


def test_rectangle_flat_axis_is_not_bounds_checked():
    """getSize zeroes the axis on which both corners agree, so rounding of the
    hit point off the plane never rejects a hit (Rectangle.cpp:16-25,32-48)."""
    rng = np.random.default_rng(5)
    n = 5000
    z = np.float32(0.1)  # not representable: p.z lands on either side of it
    q = _rect((-0.5, -0.5, z), (0.5, 0.5, z), (0.0, 0.0, -1.0))
    p = np.concatenate([rng.uniform(-0.45, 0.45, (n, 2)), np.full((n, 1), float(z))], axis=1)
    d = _unit(rng.normal(size=(n, 3)) * [1, 1, 0.2] + [0, 0, 1])
    o = p - rng.uniform(0.3, 3, n)[:, None] * d
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    hit, t = O.rectangle(rays, np.tile(q, (n, 1)))
    assert hit.all()
    pz = rays[:, 2] + t * rays[:, 5]  # FP32, as the reference computes the point
    assert (pz != z).any()  # some points are off the plane, and they still hit


def test_rectangle_parallel_and_backward_rays_miss():
    rays = np.array([[0, 0, -1, 1, 0, 0],          # parallel
                     [0, 0, -1, 1, 0, 5e-11],      # |denom| below 1e-10
                     [0, 0, -1, 1, 0, 1e-8],       # |denom| above 1e-10 but hits far outside
                     [0, 0, -1, 0, 0, -1],         # plane behind: t < 0
                     [0, 0, 1, 0, 0, 1],           # origin past the plane, moving away
                     [0, 0, 1, 0, 0, -1]], np.float32)  # from the back side, towards it: hit
    hit, t = O.rectangle(rays, np.tile(UNIT_QUAD, (len(rays), 1)))
    assert hit.tolist() == [0, 0, 0, 0, 0, 1]
    assert t[5] == 1.0


def test_rectangle_hits_match_an_fp64_plane_intersection():
    rng = np.random.default_rng(6)
    n = 20000
    axis = rng.integers(0, 3, n)
    c = rng.uniform(-1, 1, (n, 3))
    half = rng.uniform(0.05, 0.6, (n, 3))
    a, b = c - half, c + half
    nrm = np.zeros((n, 3))
    a[np.arange(n), axis] = c[np.arange(n), axis]
    b[np.arange(n), axis] = c[np.arange(n), axis]
    nrm[np.arange(n), axis] = -1.0
    rects = np.concatenate([a, b, nrm], axis=1).astype(np.float32)
    o = rng.uniform(-2, 2, (n, 3))
    d = _unit(c + rng.normal(0, 0.25, (n, 3)) - o)
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    hit, t = O.rectangle(rays, rects)
    o64, d64 = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64)
    r64 = rects.astype(np.float64)
    k = np.arange(n)
    tt = (r64[k, 3 + axis] - o64[k, axis]) / d64[k, axis]
    p = o64 + tt[:, None] * d64
    lo, hi = np.minimum(r64[:, :3], r64[:, 3:6]), np.maximum(r64[:, :3], r64[:, 3:6])
    inside = np.ones(n, bool)
    margin = np.full(n, np.inf)
    for ax in range(3):
        check = ax != axis
        inside &= ~check | ((p[:, ax] >= lo[:, ax]) & (p[:, ax] <= hi[:, ax]))
        m = np.minimum(np.abs(p[:, ax] - lo[:, ax]), np.abs(p[:, ax] - hi[:, ax]))
        margin = np.where(check, np.minimum(margin, m), margin)
    expect = inside & (tt >= 0)
    clear = (margin > 1e-4) & (np.abs(tt) > 1e-4) & (np.abs(d64[k, axis]) > 1e-3)
    assert clear.mean() > 0.98
    np.testing.assert_array_equal(hit[clear].astype(bool), expect[clear])
    assert 0.15 < hit.mean() < 0.85
    m = hit.astype(bool) & clear
    np.testing.assert_allclose(t[m], tt[m], rtol=1e-5, atol=1e-6)


# ---- GlossyBSDF::eval ------------------------------------------------------------
def _glossy(rough, metal, albedo, wi, wo):
    n = len(wi)
    return O.vec_op("neo_glossy_eval", n, np.broadcast_to(np.float32(rough), (n,)),
                    np.broadcast_to(np.float32(metal), (n,)), np.broadcast_to(np.float32(albedo), (n, 3)),
                    wi, wo, out_shape=(n, 3)).astype(np.float64)


@pytest.mark.parametrize("rough,metal", [(0.2, 0.0), (0.5, 0.3), (0.9, 1.0)])
def test_glossy_at_normal_incidence(rough, metal):
    """wi = wo = n: H = n, D = 1/(pi a^2), G = 1, F = F0 = 0.04, NdotL = 1."""
    albedo = np.array([0.8, 0.5, 0.1])
    z = np.array([[0.0, 0.0, 1.0]])
    f = _glossy(rough, metal, albedo, z, z)[0]
    a = rough * rough
    expect = albedo / np.pi * 0.96 * (1 - metal) + 0.04 / (4 * np.pi * a * a)
    # FP32: D's denominator cos^2(a^2 - 1) + 1 cancels to a^2, with a relative
    # error near 2^-24 / a^2 (3e-5 at a = 0.04)
    np.testing.assert_allclose(f, expect, rtol=max(1e-5, 3 * 2.0 ** -24 / (a * a)))


def test_glossy_is_reciprocal_and_its_diffuse_share_is_one_minus_fresnel():
    rng = np.random.default_rng(7)
    n = 4000
    wi = _unit(rng.normal(size=(n, 3)) * [1, 1, 0.5] + [0, 0, 0.6])
    wo = _unit(rng.normal(size=(n, 3)) * [1, 1, 0.5] + [0, 0, 0.6])
    keep = (wi[:, 2] > 0.05) & (wo[:, 2] > 0.05)
    wi, wo = wi[keep].astype(np.float32), wo[keep].astype(np.float32)
    albedo = np.array([0.7, 0.4, 0.2])
    for rough in (0.1, 0.4, 0.8):
        f_io = _glossy(rough, 0.25, albedo, wi, wo)
        f_oi = _glossy(rough, 0.25, albedo, wo, wi)
        # f includes NdotL (the cosine of the second argument): the BRDF itself is symmetric
        np.testing.assert_allclose(f_io / wo[:, 2:3], f_oi / wi[:, 2:3], rtol=2e-4)
        # metallic 0 minus metallic 1 leaves the diffuse share albedo/pi (1 - F) NdotL
        diff = _glossy(rough, 0.0, albedo, wi, wo) - _glossy(rough, 1.0, albedo, wi, wo)
        h = _unit(wi.astype(np.float64) + wo)
        F = 0.04 + 0.96 * (1 - np.sum(wi * h, axis=1)) ** 5
        np.testing.assert_allclose(diff, albedo / np.pi * (1 - F)[:, None] * wo[:, 2:3], rtol=2e-4, atol=1e-7)


@pytest.mark.parametrize("rough", [0.3, 0.6])
def test_glossy_specular_directional_albedo_matches_a_half_vector_quadrature(rough):
    """metallic = 1, wi = n: the outgoing-direction integral of f (which holds
    NdotL) equals the half-vector integral of D cos(theta_h) G1^2 F, with the
    reflection Jacobian 1/(4 wo.h) between them.  A dropped 4, a D evaluated at
    the wrong vector or a doubled cosine fails it."""
    a = rough * rough
    # wo-space quadrature of the oracle, nodes packed near the mirror direction
    n_mu, n_phi = 3000, 64
    edges = 1.0 - np.concatenate([[0.0], np.geomspace(1e-8, 1.0, n_mu)])[::-1]
    mu = 0.5 * (edges[1:] + edges[:-1])
    phi = (np.arange(n_phi) + 0.5) * (2 * np.pi / n_phi)
    M, P = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1 - M * M)
    wo = np.stack([s * np.cos(P), s * np.sin(P), M], -1).reshape(-1, 3).astype(np.float32)
    w = (np.diff(edges)[:, None] * np.full(n_phi, 2 * np.pi / n_phi)).reshape(-1)
    wi = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (len(wo), 1))
    f = _glossy(rough, 1.0, np.ones(3), wi, wo)[:, 0]
    lhs = np.sum(f * w)
    # half-vector space (FP64, textbook GGX and Schlick-GGX with k = a / 2)
    # only half vectors with wo = reflect(n, h) in the upper hemisphere: cos(theta_h) > 1/sqrt(2)
    c0 = np.sqrt(0.5)
    hedges = 1.0 - (1.0 - c0) * np.concatenate([[0.0], np.geomspace(1e-9, 1.0, 20000)])[::-1]
    c = 0.5 * (hedges[1:] + hedges[:-1])
    dc = np.diff(hedges)
    D = a * a / (np.pi * (c * c * (a * a - 1) + 1) ** 2)
    k = a / 2
    G1 = c / (c * (1 - k) + k)
    F = 0.04 + 0.96 * (1 - c) ** 5
    rhs = np.sum(D * c * G1 * G1 * F * dc) * 2 * np.pi
    np.testing.assert_allclose(lhs, rhs, rtol=2e-3)
    assert 0.01 < rhs < 0.05
