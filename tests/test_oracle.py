"""The oracle pinned against known answers, closed forms and brute force.

Each check is chosen so that a plausible mistake (wrong basis recurrence,
transposed mode ordering, wrong face table, sign error in the Jacobian)
fails it.  These tests run on the CPU only.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import GOLDEN, HCF_AMP, decode_fld, hex_curved_face_geom, hex_curved_face_map


# ------------------------------------------------------------------- A1 basis

def test_modA_closed_forms(oracle):
    for x in np.linspace(-1, 1, 9):
        assert oracle.modA(0, x) == pytest.approx((1 - x) / 2, abs=1e-15)
        assert oracle.modA(1, x) == pytest.approx((1 + x) / 2, abs=1e-15)
        assert oracle.modA(2, x) == pytest.approx((1 - x * x) / 4, abs=1e-15)
        # phi_3 = (1-x^2)/4 * P1^{(1,1)} = (1-x^2)/4 * 2x
        assert oracle.modA(3, x) == pytest.approx((1 - x * x) / 4 * 2 * x, abs=1e-15)
        # phi_4: P2^{(1,1)} = (3/4)(5x^2-1)
        assert oracle.modA(4, x) == pytest.approx((1 - x * x) / 4 * 0.75 * (5 * x * x - 1), abs=1e-14)


def test_modA_interior_modes_vanish_at_ends_and_are_orthogonal(oracle):
    # Bubble modes vanish at +-1; their P^{(1,1)} factors are orthogonal under
    # the weight (1-x)(1+x): int phi_p phi_q / ((1-x^2)/4)^2 * (1-x^2) dx = 0.
    xg, wg = np.polynomial.legendre.leggauss(40)
    for p in range(2, 10):
        assert oracle.modA(p, 1.0) == pytest.approx(0.0, abs=1e-15)
        assert oracle.modA(p, -1.0) == pytest.approx(0.0, abs=1e-15)
    for p in range(2, 9):
        for q in range(2, 9):
            if p == q:
                continue
            s = sum(w * oracle.modA(p, x) * oracle.modA(q, x) * 16 / (1 - x * x)
                    for x, w in zip(xg, wg))
            assert abs(s) < 1e-12, (p, q, s)


def test_modA_derivative_matches_finite_difference(oracle):
    h = 1e-6
    for p in range(0, 10):
        for x in (-0.93, -0.4, 0.0, 0.37, 0.88):
            fd = (oracle.modA(p, x + h) - oracle.modA(p, x - h)) / (2 * h)
            assert oracle.modA_d(p, x) == pytest.approx(fd, abs=1e-8)


# ------------------------------------------------- A2 ordering, separability

def test_modal_sum_is_separable_per_mode_with_r_fastest(oracle):
    n = 4
    rst = np.array([0.3, -0.55, 0.71])
    for (i, j, k) in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 3, 1), (3, 0, 2)]:
        c = np.zeros(n ** 3, np.float32)
        c[(k * n + j) * n + i] = 1.0
        v, dr, ds, dt = oracle.eval_modal(c, n, rst)
        want = oracle.modA(i, rst[0]) * oracle.modA(j, rst[1]) * oracle.modA(k, rst[2])
        assert v == pytest.approx(want, abs=1e-15)
        assert dr == pytest.approx(oracle.modA_d(i, rst[0]) * oracle.modA(j, rst[1]) *
                                   oracle.modA(k, rst[2]), abs=1e-15)
        assert dt == pytest.approx(oracle.modA(i, rst[0]) * oracle.modA(j, rst[1]) *
                                   oracle.modA_d(k, rst[2]), abs=1e-15)


# ----------------------------------------------------- Hex_CurvedFace fixture

def _hcf(oracle):
    coeffs = decode_fld(GOLDEN / "Hex_CurvedFace.fld")
    return oracle.Oracle(hex_curved_face_geom(5), coeffs.astype(np.float32), 5, 5), coeffs


def test_fld_decodes_to_known_answer():
    c = decode_fld(GOLDEN / "Hex_CurvedFace.fld")
    assert c.size == 125
    assert np.abs(c).sum() == pytest.approx(20.210065224787, abs=1e-9)


def test_hex_curved_face_geometry_closed_form(oracle):
    orc, _ = _hcf(oracle)
    rng = np.random.default_rng(1)
    for _ in range(50):
        rst = rng.uniform(-1, 1, 3)
        x, J = orc.map(0, rst)
        np.testing.assert_allclose(x, hex_curved_face_map(*rst), rtol=0, atol=1e-14)
        # Jacobian against central differences of the closed form
        h = 1e-6
        for c in range(3):
            e = np.zeros(3)
            e[c] = h
            fd = (hex_curved_face_map(*(rst + e)) - hex_curved_face_map(*(rst - e))) / (2 * h)
            np.testing.assert_allclose(J[:, c], fd, atol=1e-8)


def test_hex_curved_face_field_known_answers(oracle):
    orc, _ = _hcf(oracle)
    # float-cast coefficients (the loader casts, xml_fld_field_parser.cpp:146-152)
    assert orc.field_at(0, [0, 0, 0])[0] == pytest.approx(0.261498792600390, abs=2e-7)
    assert orc.field_at(0, [0.5, -0.25, 0.3])[0] == pytest.approx(0.403197722028196, abs=2e-7)
    for r in (-1, 1):
        for s in (-1, 1):
            for t in (-1, 1):
                want = 1.0 if (r, s, t) == (1, 1, 1) else 0.0
                assert orc.field_at(0, [r, s, t])[0] == pytest.approx(want, abs=2e-7)
    rng = np.random.default_rng(2)
    worst = 0.0
    for _ in range(500):
        rst = rng.uniform(-1, 1, 3)
        x = hex_curved_face_map(*rst)
        exact = math.sin(math.pi * x[0] / 2) * math.sin(math.pi * x[1] / 2) * math.sin(math.pi * x[2] / 2)
        worst = max(worst, abs(orc.field_at(0, rst)[0] - exact))
    assert worst < 3e-3


def test_world_gradient_matches_finite_difference_in_world_space(oracle):
    orc, _ = _hcf(oracle)
    rst = np.array([0.2, -0.3, 0.1])
    u0, gref, gw = orc.field_at(0, rst)
    x0, _ = orc.map(0, rst)
    h = 1e-6
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        rp, _ = orc.invert(0, x0 + e)
        rm, _ = orc.invert(0, x0 - e)
        fd = (orc.field_at(0, rp)[0] - orc.field_at(0, rm)[0]) / (2 * h)
        assert gw[c] == pytest.approx(fd, abs=1e-7)


def test_point_inversion_round_trip(oracle):
    orc, _ = _hcf(oracle)
    rng = np.random.default_rng(3)
    for _ in range(100):
        rst = rng.uniform(-1, 1, 3)
        x, _ = orc.map(0, rst)
        back, it = orc.invert(0, x)
        assert it > 0
        np.testing.assert_allclose(back, rst, atol=1e-12)


def test_axis_aligned_rays_hit_curved_bottom_face(oracle):
    """Closed form: o=(x0,y0,-1), d=(0,0,1) meets face 0 (t=-1) at
    z = -0.3 (1-r^2)(1-s^2), r = 2x0-1, s = 2y0-1 (0.3 as the float32
    coefficient encodes it)."""
    orc, _ = _hcf(oracle)
    for x0, y0 in [(0.5, 0.5), (0.2, 0.7), (0.9, 0.1), (0.33, 0.41)]:
        r, s = 2 * x0 - 1, 2 * y0 - 1
        ok, t, rst, _ = orc.face_hit(0, 0, [x0, y0, -1.0], [0, 0, 1.0])
        assert ok
        assert t == pytest.approx(1.0 - HCF_AMP * (1 - r * r) * (1 - s * s), abs=1e-12)
        np.testing.assert_allclose(rst, [r, s, -1.0], atol=1e-12)
        # the same ray exits through the top face t=+1 at z=1
        ok5, t5, rst5, _ = orc.face_hit(0, 5, [x0, y0, -1.0], [0, 0, 1.0])
        assert ok5 and t5 == pytest.approx(2.0, abs=1e-12)
        # and misses the side faces
        for f in (1, 2, 3, 4):
            assert not orc.face_hit(0, f, [x0, y0, -1.0], [0, 0, 1.0])[0]


def test_oblique_ray_face_hit_satisfies_map_equation(oracle):
    orc, _ = _hcf(oracle)
    o = np.array([1.7, -0.6, 0.4])
    d = np.array([-1.0, 0.9, 0.05])
    d /= np.linalg.norm(d)
    best = None
    for f in range(6):
        ok, t, rst, _ = orc.face_hit(0, f, o, d)
        if ok:
            x, _ = orc.map(0, rst)
            np.testing.assert_allclose(x, o + t * d, atol=1e-11)
            assert np.max(np.abs(rst)) <= 1 + 1e-4 + 1e-12
            best = t if best is None else min(best, t)
    assert best is not None


def test_iso_on_linear_field_is_plane(oracle):
    """Field u = z (coefficients of the z map component): the isosurface
    u = c is the curved-map preimage of z = c, so along an axis-aligned ray
    the root is at z = c exactly."""
    geom = hex_curved_face_geom(5)
    orc = oracle.Oracle(geom, geom[2].copy(), 5, 5)
    o = np.array([0.3, 0.6, -1.0])
    d = np.array([0.0, 0.0, 1.0])
    ok0, ta, _, _ = orc.face_hit(0, 0, o, d)
    ok5, tb, _, _ = orc.face_hit(0, 5, o, d)
    for c in (0.1, 0.45, 0.8):
        ok, t, rst, u = orc.iso_segment(0, o, d, ta, tb, c)
        assert ok
        assert t == pytest.approx(1.0 + c, abs=1e-12)
        assert u == pytest.approx(c, abs=1e-12)


def _field_of_t(oracle, f, nf):
    """Modal coefficients (nf^3) of u(r, s, t) = f(t), f a polynomial of
    degree < nf: the t-modes solve f at nf points; constant in r and s through
    phi_0 + phi_1 = 1."""
    ts = np.linspace(-0.9, 0.9, nf)
    A = np.array([[oracle.lib().orc_modA(k, x) for k in range(nf)] for x in ts])
    ck = np.linalg.solve(A, f(ts))
    c = np.zeros((nf, nf, nf))
    for k in range(nf):
        c[k, :2, :2] = ck[k]
    return c.ravel()


@pytest.mark.parametrize("roots,tie", [((0.06,), False), ((0.03, 0.06, 0.09), True)])
def test_iso_several_crossings_in_one_sample_interval_are_ties(oracle, roots, tie):
    """Straight unit cube, u = 1e3 prod (t - t_i) in the reference t (z =
    (t + 1) / 2): with all roots inside the 16-sample scan's interval
    z in [0.5, 0.5625] the scan sees one sign change, but which crossing an
    endgame converges to is not determined, so the oracle marks it a tie
    (margin 0); a single root is not a tie.  The root found is one of them."""
    geom = hex_curved_face_geom(5)
    geom[2, (0 * 5 + 2) * 5 + 2] = 0.0          # flat bottom face
    nf = 4

    def f(x):
        return 1e3 * np.prod([x - r for r in roots], axis=0)
    field = _field_of_t(oracle, f, nf)
    orc = oracle.Oracle(geom, field.astype(np.float32), 5, nf)
    o = np.array([0.3, 0.6, -1.0])
    d = np.array([0.0, 0.0, 1.0])
    ok, t, rst, u = orc.iso_segment(0, o, d, 1.0, 2.0, 0.0)
    assert ok
    z = t - 1.0
    assert min(abs((2 * z - 1) - r) for r in roots) < 1e-4   # float32 coefficients move the roots
    rays = np.array([[*o, *d]], np.float32).T
    _, iso = orc.trace(rays, [0.0])
    assert (iso["minabs"][0][0] == 0.0) == tie


def test_camera_rays_closed_form(oracle):
    cam = np.array([0.5 + 2.5, 0.5 - 2.5, 0.5 - 2.0, 0.5, 0.5, 0.5, 0, 1, 0, 20.0])
    W, H = 64, 36
    rays = oracle.camera_rays(cam, W, H)
    look = cam[3:6] - cam[0:3]
    look /= np.linalg.norm(look)
    d = rays[3:6]
    np.testing.assert_allclose(np.linalg.norm(d, axis=0), 1.0, atol=1e-14)
    # the four pixels around the image centre straddle the look direction
    c = [(H // 2) * W + W // 2, (H // 2 - 1) * W + W // 2 - 1]
    mid = d[:, c].mean(axis=1)
    mid /= np.linalg.norm(mid)
    np.testing.assert_allclose(mid, look, atol=1e-6)
    # vertical field of view: angle between bottom-row and top-row centre rays
    top = d[:, (H - 1) * W + W // 2]
    bot = d[:, W // 2]
    ang = math.degrees(math.acos(np.clip(top @ bot, -1, 1)))
    assert ang == pytest.approx(20.0 * (H - 1) / H, rel=2e-3)


def _bowl_crossings(o, d, amp=HCF_AMP, n=20001, t1=3.0):
    """Brute force, independent of any Newton: crossings of the ray with the
    fixture's bottom face z = -amp (1 - r^2)(1 - s^2), x = (r+1)/2,
    y = (s+1)/2 over the unit square, by dense sign changes of the ray's
    height above the surface and FP64 bisection."""
    o, d = np.asarray(o, float), np.asarray(d, float)

    def h(t):
        p = o + t * d
        r, s = 2 * p[0] - 1, 2 * p[1] - 1
        return p[2] + amp * (1 - r * r) * (1 - s * s), abs(r) <= 1 and abs(s) <= 1
    ts = np.linspace(0.0, t1, n)
    roots = []
    prev = h(ts[0])
    for t in ts[1:]:
        cur = h(t)
        if prev[0] * cur[0] <= 0 and (prev[1] or cur[1]):
            a, b = t - (ts[1] - ts[0]), t
            for _ in range(200):
                m = 0.5 * (a + b)
                if h(a)[0] * h(m)[0] <= 0:
                    b = m
                else:
                    a = m
            if h(0.5 * (a + b))[1]:
                roots.append(0.5 * (a + b))
        prev = cur
    return roots


@pytest.mark.parametrize("z0,y0,dy,dz", [(-0.05, 0.5, 0.0, 0.0), (-0.12, 0.43, 0.07, 0.01),
                                         (-0.2, 0.6, -0.05, -0.02), (-0.27, 0.5, 0.0, 0.015)])
def test_face_crossed_twice_nearest_root_matches_brute_force(oracle, z0, y0, dy, dz):
    """A nearly horizontal ray below z = 0 enters the bowl of the curved
    bottom face and leaves it again: the face is crossed twice.  The oracle's
    face root must be the nearer crossing that dense sampling finds."""
    orc, _ = _hcf(oracle)
    o = np.array([-0.5, y0, z0], np.float32).astype(float)
    d = np.array([1.0, dy, dz], np.float32).astype(float)
    roots = _bowl_crossings(o, d)
    assert len(roots) == 2, roots
    ok, t, rst, _ = orc.face_hit(0, 0, o, d)
    assert ok
    assert t == pytest.approx(roots[0], abs=1e-10)
    x, _ = orc.map(0, rst)
    np.testing.assert_allclose(x, o + t * d, atol=1e-11)
