"""Host-side parts of the library, on the CPU: the C ABI exports, the
structured curved-hex generator against its closed form, exact Bernstein
boxes against dense sampling, and BVH invariants.  No GPU call is made."""
from __future__ import annotations

import ctypes as C
import itertools
import math
import re

import numpy as np
import pytest

from conftest import ROOT


def test_library_exports_every_declared_symbol(femrt):
    header = (ROOT / "include" / "femrt.h").read_text()
    declared = set(re.findall(r"FEMRT_API\s+[\w\s\*]+?\b(femrt_\w+)\s*\(", header))
    assert declared, "no declarations parsed"
    assert declared == set(femrt.exported_symbols())
    lib = C.CDLL(str(femrt.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert femrt.lib().femrt_version() >= 100


def _lagrange(nodes, x):
    out = np.ones(len(nodes))
    for a in range(len(nodes)):
        for b in range(len(nodes)):
            if a != b:
                out[a] *= (x - nodes[b]) / (nodes[a] - nodes[b])
    return out


FACES = [  # fix axis, side (0:-1, 1:+1), u axis, v axis (Nektar order)
    (2, 0, 0, 1), (1, 0, 0, 2), (0, 1, 1, 2), (1, 1, 0, 2), (0, 0, 1, 2), (2, 1, 0, 1)]


def closed_form_map(N, q, amp, ijk, rst):
    """Trilinear map plus, per face, the degree-q interpolant of the evenly
    spaced face-node displacements (sinusoid amp sin(2 pi a) sin(2 pi b) of
    the node's in-plane world coordinates, zero on the face boundary) times
    the linear weight (1 -+ w)/2 in the face-normal coordinate."""
    h = 1.0 / N
    x = np.array([(ijk[c] + 0.5 * (rst[c] + 1)) * h for c in range(3)])
    even = np.array([-1 + 2 * a / q for a in range(q + 1)])
    for fix, side, ua, va in FACES:
        lu = _lagrange(even, rst[ua])
        lv = _lagrange(even, rst[va])
        D = 0.0
        for a in range(1, q):
            for b in range(1, q):
                w = np.zeros(3)
                w[fix] = (ijk[fix] + side) * h
                w[ua] = (ijk[ua] + a / q) * h
                w[va] = (ijk[va] + b / q) * h
                a0, a1 = [c for c in range(3) if c != fix]
                disp = amp * math.sin(2 * math.pi * w[a0]) * math.sin(2 * math.pi * w[a1])
                D += disp * lu[a] * lv[b]
        wn = 0.5 * (1 + rst[fix]) if side else 0.5 * (1 - rst[fix])
        x[fix] += D * wn
    return x


@pytest.mark.parametrize("N,q,p,amp", [(4, 3, 4, 0.05), (2, 4, 6, 0.03)])
def test_generator_matches_closed_form(femrt, oracle, N, q, p, amp):
    hm = femrt.gen_structured_hex(N, q, p, pert=1, amp=amp, seed=7)
    orc = oracle.Oracle.from_mesh(hm)
    rng = np.random.default_rng(4)
    worst = 0.0
    for _ in range(200):
        ijk = rng.integers(0, N, 3)
        e = int((ijk[2] * N + ijk[1]) * N + ijk[0])
        rst = rng.uniform(-1, 1, 3)
        x, _ = orc.map(e, rst)
        worst = max(worst, np.max(np.abs(x - closed_form_map(N, q, amp, ijk, rst))))
    assert worst < 2e-6  # FP32 coefficient rounding


def test_generator_corners_adjacency_and_conformity(femrt, oracle):
    N = 3
    hm = femrt.gen_structured_hex(N, 4, 3, pert=2, amp=0.08, seed=11)
    orc = oracle.Oracle.from_mesh(hm)
    corners = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
               (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    rng = np.random.default_rng(5)
    opp = [5, 3, 4, 1, 2, 0]
    for e in range(hm.n_elem):
        for c, rst in enumerate(corners):  # main.cu:93-121 vertex reproduction
            x, _ = orc.map(e, rst)
            np.testing.assert_allclose(x, hm.vertices[e, c], atol=2e-7)
        for f in range(6):
            ne, nfc = hm.adjacency[e, f]
            if ne < 0:
                continue
            assert nfc == opp[f] and hm.adjacency[ne, nfc, 0] == e
            fix, side, ua, va = FACES[f]
            for _ in range(5):
                u, v = rng.uniform(-1, 1, 2)
                a = np.zeros(3)
                a[fix] = 1 if side else -1
                a[ua], a[va] = u, v
                fix2, side2, ua2, va2 = FACES[nfc]
                b = np.zeros(3)
                b[fix2] = 1 if side2 else -1
                b[ua2], b[va2] = u, v
                xa, _ = orc.map(e, a)
                xb, _ = orc.map(ne, b)
                np.testing.assert_allclose(xa, xb, atol=1e-6)


@pytest.mark.parametrize("N,q,pert,amp,seed,n_check", [(4, 3, 1, 0.05, 1, 64), (6, 4, 2, 0.08, 5, 216),
                                                      (32, 4, 2, 0.08, 20240611, 150),
                                                      (64, 4, 3, 0.15, 404, 150)])
def test_generator_elements_are_valid(femrt, oracle, N, q, pert, amp, seed, n_check):
    """Every checked element has a positive Jacobian determinant on a 5^3 grid
    (the map is one-to-one, so point inversion and face roots are defined)."""
    hm = femrt.gen_structured_hex(N, q, 2, pert=pert, amp=amp, seed=seed)
    orc = oracle.Oracle.from_mesh(hm)
    rng = np.random.default_rng(0)
    g = np.linspace(-1, 1, 5)
    els = range(hm.n_elem) if hm.n_elem <= n_check else rng.choice(hm.n_elem, n_check, replace=False)
    worst = min(np.linalg.det(orc.map(e, np.array(p))[1]) * (2.0 * N) ** 3
                for e in els for p in itertools.product(g, g, g))
    assert worst > 0.15


def test_generator_field_is_seeded_and_decays(femrt):
    a = femrt.gen_structured_hex(2, 3, 6, seed=3)
    b = femrt.gen_structured_hex(2, 3, 6, seed=3)
    c = femrt.gen_structured_hex(2, 3, 6, seed=4)
    np.testing.assert_array_equal(a.field, b.field)
    assert not np.array_equal(a.field, c.field)
    nf = 7
    deg = np.add.outer(np.add.outer(np.arange(nf), np.arange(nf)), np.arange(nf)).ravel()
    scaled = a.field * (1.0 + deg)[None, :] ** 2  # (k,j,i) symmetric sum of indices
    assert abs(scaled.mean()) < 0.1 and 0.8 < scaled.std() < 1.2


@pytest.mark.parametrize("pert,amp,q", [(1, 0.05, 3), (2, 0.08, 4)])
def test_bernstein_boxes_contain_dense_samples(femrt, oracle, pert, amp, q):
    N = 3
    hm = femrt.gen_structured_hex(N, q, 3, pert=pert, amp=amp, seed=9)
    boxes = femrt.host_boxes(hm)
    orc = oracle.Oracle.from_mesh(hm)
    g = np.linspace(-1, 1, 9)
    rng = np.random.default_rng(6)
    for e in range(hm.n_elem):
        lo, hi = boxes[e, 0, :3], boxes[e, 0, 3:]
        pts = [orc.map(e, rst)[0] for rst in itertools.product(g, g, g)]
        pts += [orc.map(e, rng.uniform(-1, 1, 3))[0] for _ in range(100)]
        pts = np.array(pts)
        assert np.all(pts >= lo) and np.all(pts <= hi)
        ext = pts.max(0) - pts.min(0)
        assert np.all((hi - lo) <= 1.6 * ext + 1e-6)  # tight enough to be useful
        for f, (fix, side, ua, va) in enumerate(FACES):
            flo, fhi = boxes[e, 1 + f, :3], boxes[e, 1 + f, 3:]
            for u, v in itertools.product(g, g):
                rst = np.zeros(3)
                rst[fix] = 1 if side else -1
                rst[ua], rst[va] = u, v
                x, _ = orc.map(e, rst)
                assert np.all(x >= flo) and np.all(x <= fhi), (e, f)
                # the sub-patch boxes that contain (u, v) contain the point
                for a in range(3):
                    for b in range(3):
                        if not (-1 + 2 * a / 3 - 1e-12 <= u <= -1 + 2 * (a + 1) / 3 + 1e-12 and
                                -1 + 2 * b / 3 - 1e-12 <= v <= -1 + 2 * (b + 1) / 3 + 1e-12):
                            continue
                        sb = boxes[e, 7 + 9 * f + 3 * b + a]
                        assert np.all(x >= sb[:3]) and np.all(x <= sb[3:]), (e, f, a, b)
            # face roots are accepted up to |u|, |v| <= 1 + 1e-4: the band just
            # outside the face lies in the element, face and edge sub-patch boxes
            acc = 1 + 1e-4
            for u, v in [(acc, 0.3), (-acc, -0.7), (0.1, acc), (-0.4, -acc), (acc, acc), (-acc, acc)]:
                rst = np.zeros(3)
                rst[fix] = 1 if side else -1
                rst[ua], rst[va] = u, v
                x, _ = orc.map(e, rst)
                a = min(2, int((np.clip(u, -1, 1) + 1) * 1.5))
                b = min(2, int((np.clip(v, -1, 1) + 1) * 1.5))
                for bx in (boxes[e, 0], boxes[e, 1 + f], boxes[e, 7 + 9 * f + 3 * b + a]):
                    assert np.all(x >= bx[:3]) and np.all(x <= bx[3:]), (e, f, u, v)


def test_bvh_invariants(femrt):
    hm = femrt.gen_structured_hex(6, 3, 2, pert=2, amp=0.08, seed=1)
    boxes = femrt.host_boxes(hm)
    nodes, prims = femrt.host_build_bvh(boxes)
    assert sorted(prims.tolist()) == list(range(hm.n_elem))
    ab = nodes.view(np.int32)
    seen = np.zeros(hm.n_elem, int)

    def visit(i, depth):
        lo, hi = nodes[i, :3], nodes[i, 4:7]
        a, b = ab[i, 3], ab[i, 7]
        if b < 0:
            assert 1 <= -b <= 4
            for s in range(a, a - b):
                e = prims[s]
                seen[e] += 1
                assert np.all(boxes[e, 0, :3] >= lo) and np.all(boxes[e, 0, 3:] <= hi)
            return depth
        for c in (a, b):
            assert np.all(nodes[c, :3] >= lo) and np.all(nodes[c, 4:7] <= hi)
        return max(visit(a, depth + 1), visit(b, depth + 1))

    depth = visit(0, 0)
    assert np.all(seen == 1)
    assert depth < 40


@pytest.mark.parametrize("p", [2, 4, 6])
def test_field_bounds_contain_dense_samples(femrt, oracle, p):
    """The per-element field bounds the walk uses to skip scans contain u at
    a dense grid and random points of every element, and are not vacuous."""
    hm = femrt.gen_structured_hex(2, 3, p, pert=1, amp=0.05, seed=4)
    fr = femrt.host_field_range(hm)
    nf = hm.nf
    field = hm.field.reshape(hm.n_elem, -1)
    g = np.linspace(-1, 1, 9)
    rng = np.random.default_rng(3)
    for e in range(hm.n_elem):
        pts = list(itertools.product(g, g, g)) + [tuple(rng.uniform(-1, 1, 3)) for _ in range(100)]
        u = np.array([oracle.eval_modal(field[e], nf, np.array(r))[0] for r in pts])
        assert fr[e, 0] <= u.min() and u.max() <= fr[e, 1], e
        assert fr[e, 1] - fr[e, 0] <= 1.5 * (u.max() - u.min()) + 1e-4, e


def test_field_bounds_closed_forms(femrt):
    """Bernstein bounds are exact for a constant and for a linear field
    (u = 0.7 from the vertex modes; u = r from phi_1 - phi_0), up to the pad
    of 1e-5 sum |c|."""
    nf = 5
    c = np.zeros((2, nf, nf, nf), np.float32)   # [elem][k][j][i]
    c[0, :2, :2, :2] = 0.7
    c[1, :2, :2, 1] = 1.0
    c[1, :2, :2, 0] = -1.0
    hm = femrt.HexMesh(np.zeros((2, 24), np.float32), np.zeros((2, 3 * 64), np.float32),
                       c.reshape(2, -1), np.full((2, 12), -1, np.int32), 4, nf)
    fr = femrt.host_field_range(hm)
    np.testing.assert_allclose(fr[0], [0.7, 0.7], atol=1e-4)
    assert fr[0, 0] < 0.7 < fr[0, 1]
    np.testing.assert_allclose(fr[1], [-1.0, 1.0], atol=1e-4)
    assert fr[1, 0] < -1.0 and fr[1, 1] > 1.0


@pytest.mark.parametrize("p", [2, 4, 6, 8])
def test_field_lipschitz_bounds_the_gradient(femrt, oracle, p):
    """The bound the walk's early sign acceptance uses: |u(a) - u(b)| <=
    L |a - b|_inf on the reference cube, i.e. L >= |du/dr| + |du/ds| +
    |du/dt| everywhere (checked on a dense grid and random points), and L is
    within 3 times that maximum (the half-cube Bernstein bound)."""
    hm = femrt.gen_structured_hex(2, 3, p, pert=1, amp=0.05, seed=8)
    L = femrt.host_field_lipschitz(hm)
    nf = hm.nf
    field = hm.field.reshape(hm.n_elem, -1)
    g = np.linspace(-1, 1, 9)
    rng = np.random.default_rng(5)
    for e in range(hm.n_elem):
        pts = list(itertools.product(g, g, g)) + [tuple(rng.uniform(-1, 1, 3)) for _ in range(200)]
        g1 = max(np.abs(oracle.eval_modal(field[e], nf, np.array(r))[1:]).sum() for r in pts)
        assert g1 <= L[e], (e, g1, L[e])
        assert L[e] <= 3.0 * g1, (e, g1, L[e])
        # pairs of points: the Lipschitz statement itself
        for _ in range(50):
            a, b = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
            du = abs(oracle.eval_modal(field[e], nf, a)[0] - oracle.eval_modal(field[e], nf, b)[0])
            assert du <= L[e] * np.abs(a - b).max() * (1 + 1e-6) + 1e-12


def test_field_lipschitz_closed_forms(femrt):
    """u = 0.7 (vertex modes): L = 0; u = r (phi_1 - phi_0): L = 1.05 (the
    gradient is exactly (1, 0, 0), its Bernstein coefficients are all 1, plus
    the 5% margin); u = r + 2 s: L = 3.15."""
    nf = 5
    c = np.zeros((3, nf, nf, nf), np.float32)   # [elem][k][j][i]
    c[0, :2, :2, :2] = 0.7
    c[1, :2, :2, 1] = 1.0
    c[1, :2, :2, 0] = -1.0
    c[2, :2, :2, 1] += 1.0
    c[2, :2, :2, 0] -= 1.0
    c[2, :2, 1, :2] += 2.0
    c[2, :2, 0, :2] -= 2.0
    hm = femrt.HexMesh(np.zeros((3, 24), np.float32), np.zeros((3, 3 * 64), np.float32),
                       c.reshape(3, -1), np.full((3, 12), -1, np.int32), 4, nf)
    L = femrt.host_field_lipschitz(hm)
    np.testing.assert_allclose(L, [0.0, 1.05, 3.15], atol=1e-5)


@pytest.mark.parametrize("N,q,pert,amp,found", [(4, 3, 1, 0.05, 0.0), (6, 4, 2, 0.08, 0.95),
                                                 (8, 4, 3, 0.15, 0.95)])
def test_geom_jinv_bound_holds(femrt, oracle, N, q, pert, amp, found):
    """The bound K of |J^-1|_inf (max row sum) that certifies skipped scan
    samples: where one is found (K > 0), K >= |J^-1|_inf on a dense grid and
    random points of the element, and within 6x of the sampled maximum; it
    is found for most elements of the benchmark-like meshes (the strongly
    sheared 4^3 mesh with sinusoidal faces of 0.2 h gets none, and its walk
    simply skips nothing).  On the unit cube split in N^3 straight hexes
    J^-1 = (2N) I exactly, so K is then 2N (+5%)."""
    hm = femrt.gen_structured_hex(N, q, 2, pert=pert, amp=amp, seed=12)
    K = femrt.host_geom_jinv_bound(hm)
    assert np.mean(K > 0) >= found
    orc = oracle.Oracle.from_mesh(hm)
    rng = np.random.default_rng(2)
    g = np.linspace(-1, 1, 6)
    for e in rng.choice(hm.n_elem, min(hm.n_elem, 24), replace=False):
        if K[e] == 0:
            continue
        pts = list(itertools.product(g, g, g)) + [tuple(rng.uniform(-1, 1, 3)) for _ in range(100)]
        worst = max(np.abs(np.linalg.inv(orc.map(e, np.array(r))[1])).sum(1).max() for r in pts)
        assert worst <= K[e], (e, worst, K[e])
        assert K[e] <= 6.0 * worst, (e, worst, K[e])
    flat = femrt.gen_structured_hex(3, 3, 2, pert=0, seed=1)
    np.testing.assert_allclose(femrt.host_geom_jinv_bound(flat), 6.0 * 1.05, rtol=1e-5)


def test_field_waves_is_continuous_seeded_and_unit_rms(femrt, oracle):
    """The multi-frequency field interpolates one smooth function at each
    element's GLL nodes: both sides of every checked interior face give the
    same u (the face trace of the interpolant depends on the face nodes
    only), it is seeded, and its RMS over the cube is about 1."""
    hm = femrt.gen_field_waves(femrt.gen_structured_hex(8, 4, 2, pert=2, amp=0.15, seed=3), 8,
                               n_waves=8, seed=5)
    assert hm.nf == 9 and hm.field.shape == (512, 729)
    again = femrt.gen_field_waves(hm, 8, n_waves=8, seed=5)
    other = femrt.gen_field_waves(hm, 8, n_waves=8, seed=6)
    assert np.array_equal(hm.field, again.field) and not np.array_equal(hm.field, other.field)
    orc = oracle.Oracle.from_mesh(hm)
    rng = np.random.default_rng(0)
    # faces 2 (r = +1) and 4 (r = -1) of x-neighbours, 1 (s = -1) / 3 (s = +1)
    pairs = [(2, 0, 1.0, -1.0), (3, 1, 1.0, -1.0)]
    n_checked = 0
    for _ in range(150):
        e = int(rng.integers(0, hm.n_elem))
        for f, axis, side_a, side_b in pairs:
            ne = hm.adjacency[e, f, 0]
            if ne < 0:
                continue
            a, b = rng.uniform(-1, 1, 2)
            pa = np.insert(np.array([a, b]), axis, side_a)
            pb = np.insert(np.array([a, b]), axis, side_b)
            xa = orc.map(e, pa)[0]
            xb = orc.map(ne, pb)[0]
            np.testing.assert_allclose(xa, xb, atol=1e-6)
            assert abs(orc.field_at(e, pa)[0] - orc.field_at(ne, pb)[0]) < 1e-5
            n_checked += 1
    assert n_checked > 100
    us = np.array([orc.field_at(int(rng.integers(0, hm.n_elem)), rng.uniform(-1, 1, 3))[0]
                   for _ in range(3000)])
    assert 0.7 < np.sqrt(np.mean(us ** 2)) < 1.3
