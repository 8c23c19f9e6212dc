"""Curved quadrilateral surface elements (the reference's QuadElementData,
fem/QuadElement.cu; scene setup raytracer.h:213-257): the FP64 oracle pinned
by closed forms, and the GPU first hits against it.

The oracle's quad map is the 2D Modified_A expansion with mode (i, j) at
j * n + i (EvaluateModal2D, fem/QuadElement.cu:22-37) and its intersection
the nearest Newton root on (r, s, t) over every quad (IntersectRayWithQuad,
fem/QuadElement.cu:467-527, seeds by sampling).  The pins below fix the
layout and the root independently of that restatement: a flat quad is a
plane with a bilinear parameterisation, and a single mode-(2, 2) coefficient
16 a h raises the quad by a h (1 - r^2)(1 - s^2), phi_2(x) = (1 - x^2) / 4
being the textbook Modified_A bubble."""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle.oracle as orc


def _one_quad(x0=0.0, y0=0.0, h=1.0, modes=None, field=None, ng=3, nf=3):
    fm = pytest.importorskip("importlib").import_module("test-_amd")
    v = np.array([[x0, y0, 0], [x0 + h, y0, 0], [x0, y0 + h, 0], [x0 + h, y0 + h, 0]], np.float32)
    g = np.zeros((1, 3, ng * ng), np.float32)
    for q in range(4):
        i, j = q & 1, q >> 1
        g[0, :, j * ng + i] = v[q]
    for (c, i, j), val in (modes or {}).items():
        g[0, c, j * ng + i] = val
    f = np.zeros((1, nf * nf), np.float32) if field is None else np.asarray(field, np.float32).reshape(1, -1)
    return fm.QuadMesh(v[None], g, f, ng, nf)


def _rays(o, d):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    return np.ascontiguousarray(np.concatenate([o.T, d.T]), np.float32)


def test_modified_a_bubble_closed_form(oracle):
    for x in np.linspace(-1, 1, 9):
        assert orc.modA(2, x) == pytest.approx((1 - x * x) / 4, abs=1e-15)
        assert orc.modA(0, x) == pytest.approx((1 - x) / 2, abs=1e-15)
        assert orc.modA(1, x) == pytest.approx((1 + x) / 2, abs=1e-15)


def test_flat_quad_is_a_plane(oracle):
    qm = _one_quad(0.5, -0.25, 2.0)
    Q = orc.QuadOracle.from_mesh(qm)
    rng = np.random.default_rng(1)
    px = rng.uniform(0.5, 2.5, 50)
    py = rng.uniform(-0.25, 1.75, 50)
    o = np.stack([px + 0.3, py - 0.2, np.full(50, 2.0)], 1)
    d = np.tile([-0.15, 0.1, -1.0], (50, 1))
    og = Q.trace(_rays(o, d))
    o32, d32 = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    t = -o32[:, 2] / d32[:, 2]
    x = o32 + t[:, None] * d32
    assert np.all(og["elem"] == 0) and np.all(og["face"] == 0)
    np.testing.assert_allclose(og["t"], t, rtol=1e-12)
    # bilinear map of an axis-aligned square: r = 2 (x - x0) / h - 1
    np.testing.assert_allclose(og["rst"][0], (x[:, 0] - 0.5) - 1, atol=1e-12)
    np.testing.assert_allclose(og["rst"][1], (x[:, 1] + 0.25) - 1, atol=1e-12)
    assert np.all(og["rst"][2] == -1)


def test_bump_height_and_mode_order(oracle):
    a = 0.3
    # bubble in z, and a mode (2, 0) term in y that a transposed layout would
    # place along the other axis
    qm = _one_quad(0.0, 0.0, 1.0, modes={(2, 2, 2): 16 * a, (1, 2, 0): 0.2})
    Q = orc.QuadOracle.from_mesh(qm)
    for r, s in ((0.0, 0.0), (0.3, -0.7), (-0.9, 0.5)):
        x, xr, xs = Q.map(0, r, s)
        assert x[2] == pytest.approx(a * (1 - r * r) * (1 - s * s), abs=1e-6)
        # y = (s + 1)/2 + 0.2 phi_2(r) phi_0(s)
        assert x[1] == pytest.approx((s + 1) / 2 + 0.2 * (1 - r * r) / 4 * (1 - s) / 2, abs=1e-6)
        assert xr[1] == pytest.approx(0.2 * (-2 * r) / 4 * (1 - s) / 2, abs=1e-6)
    # vertical rays onto the bump: t = z0 - z(x, y), the (r, s) of the hit
    # solving the map, the field the quad's coefficients at (r, s)
    field = np.arange(9, dtype=np.float32) / 9 - 0.4
    qm = _one_quad(0.0, 0.0, 1.0, modes={(2, 2, 2): 16 * a}, field=field)
    Q = orc.QuadOracle.from_mesh(qm)
    rs = np.array([[0.1, 0.2], [-0.5, 0.8], [0.95, -0.95]])
    o = np.stack([(rs[:, 0] + 1) / 2, (rs[:, 1] + 1) / 2, np.full(3, 3.0)], 1)
    og = Q.trace(_rays(o, np.tile([0, 0, -1.0], (3, 1))))
    o32 = o.astype(np.float32).astype(np.float64)
    r, s = 2 * o32[:, 0] - 1, 2 * o32[:, 1] - 1
    a32 = float(np.float32(16 * a)) / 16     # the coefficient as stored
    np.testing.assert_allclose(og["t"], 3.0 - a32 * (1 - r * r) * (1 - s * s), rtol=1e-12)
    np.testing.assert_allclose(og["rst"][0], r, atol=1e-12)
    np.testing.assert_allclose(og["rst"][1], s, atol=1e-12)
    phi = lambda x: np.array([(1 - x) / 2, (1 + x) / 2, (1 - x * x) / 4])  # noqa: E731
    for k in range(3):
        u = sum(float(field[j * 3 + i]) * phi(r[k])[i] * phi(s[k])[j] for i in range(3) for j in range(3))
        assert og["u"][k] == pytest.approx(u, abs=1e-12)


def test_nearest_of_stacked_quads_and_edge_miss(oracle):
    fm = pytest.importorskip("importlib").import_module("test-_amd")
    a, b = _one_quad(0, 0, 1), _one_quad(0, 0, 1)
    for q in range(4):          # the same square lifted to z = 0.5
        i, j = q & 1, q >> 1
        b.geom[0, 2, j * 3 + i] = 0.5
        b.vertices[0, q, 2] = 0.5
    qm = fm.QuadMesh(np.concatenate([a.vertices, b.vertices]), np.concatenate([a.geom, b.geom]),
                     np.zeros((2, 9)), 3, 3)
    Q = orc.QuadOracle.from_mesh(qm)
    og = Q.trace(_rays([[0.5, 0.5, 2.0], [0.5, 0.5, -2.0], [1.00002, 0.5, 2.0], [1.01, 0.5, 2.0]],
                       [[0, 0, -1.0], [0, 0, 1.0], [0, 0, -1.0], [0, 0, -1.0]]))
    assert list(og["elem"][:2]) == [1, 0]
    assert og["t"][0] == pytest.approx(1.5) and og["t2"][0] == pytest.approx(2.0)
    assert og["t"][1] == pytest.approx(2.0) and og["t2"][1] == pytest.approx(2.5)
    # 2e-5 outside the edge (r = 1.00004): inside the acceptance bound, a
    # tie by its margin; 1e-2 outside: a miss
    assert og["elem"][2] == 1 and og["margin"][2] < 1e-4
    assert og["elem"][3] == -1


def test_bumpy_scene_is_a_height_field(femrt, oracle):
    """The bumpy scene is the graph of z(x, y) over the unit square, so every
    vertical ray inside it hits once, at z(x, y): small quads (Jacobian
    1/32) converge like large ones."""
    qm = _scene(femrt)
    Q = orc.QuadOracle.from_mesh(qm)
    rng = np.random.default_rng(0)
    n = 2000
    x, y = rng.uniform(0.001, 0.999, n), rng.uniform(0.001, 0.999, n)
    rays = np.stack([x, y, np.full(n, 5.0), np.zeros(n), np.zeros(n), -np.ones(n)]).astype(np.float32)
    og = Q.trace(rays)
    assert np.all(og["elem"] >= 0)
    # the hit's quad is the grid cell of (x, y), and the map puts it there
    x32, y32 = rays[0].astype(np.float64), rays[1].astype(np.float64)
    cell = np.floor(y32 * 16).astype(int) * 16 + np.floor(x32 * 16).astype(int)
    inside = (np.abs(x32 * 16 - np.round(x32 * 16)) > 1e-6) & (np.abs(y32 * 16 - np.round(y32 * 16)) > 1e-6)
    assert np.all(og["elem"][inside] == cell[inside])
    for k in range(0, n, 97):
        X, _, _ = Q.map(og["elem"][k], og["rst"][0][k], og["rst"][1][k])
        assert X[0] == pytest.approx(x32[k], abs=1e-12) and X[1] == pytest.approx(y32[k], abs=1e-12)
        assert 5.0 - og["t"][k] == pytest.approx(X[2], abs=1e-12)


def _scene(femrt, n=16, seed=5):
    qm = femrt.gen_bumpy_quads(n, amp=0.35, seed=seed, tilt=0.25)
    return qm


@pytest.mark.gpu
@pytest.mark.parametrize("builder", ["host_sah", "gpu_lbvh"])
def test_quads_first_hit_parity(femrt, oracle, gpu_ctx, builder):
    import oracle.parity as parity
    qm = _scene(femrt)
    mesh = femrt.Mesh.from_quads(gpu_ctx, qm, build_bvh=False)
    mesh.build_bvh(femrt.BVH_HOST_SAH if builder == "host_sah" else femrt.BVH_GPU_LBVH)
    W, H = 160, 120
    fr = femrt.Frame(gpu_ctx, W * H)
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 0.3])
    fr.camera_rays(cam, W, H)
    fr.trace(mesh, femrt.trace_params([], iso_enable=False))
    gh = fr.hits()
    rays = fr.get_rays()
    og = orc.QuadOracle.from_mesh(qm).trace(rays)
    rep = parity.compare(gh, None, og, None, [])
    print(json.dumps({"quads": qm.n_elem, "builder": builder, "report": rep}), flush=True)
    assert rep["oracle_hits"] > 0.2 * W * H
    assert rep["ok"]
    # the hit is face 0 at t = -1 of the embedding
    h = gh["elem"] >= 0
    assert np.all(gh["face"][h] == 0) and np.all(gh["rst"][2][h] == -1)
    fr.close()
    mesh.close()


@pytest.mark.gpu
def test_quads_reject_isosurfaces(femrt, gpu_ctx):
    mesh = femrt.Mesh.from_quads(gpu_ctx, _scene(femrt, 4))
    fr = femrt.Frame(gpu_ctx, 64)
    fr.camera_rays(femrt.auto_camera([0, 0, 0], [1, 1, 0.3]), 8, 8)
    with pytest.raises(RuntimeError, match="surface mesh"):
        fr.trace(mesh, femrt.trace_params([0.1]))
    fr.close()
    mesh.close()


@pytest.mark.gpu
def test_cpp_scene_add_quad_render(femrt, tmp_path):
    """Scene::add_quad through the command-line renderer: the quads of a
    bumpy scene, rendered from the automatic camera, against the oracle on
    the renderer's own rays."""
    import subprocess

    import oracle.parity as parity
    from conftest import ROOT
    from test_dropin import AOV
    qm = _scene(femrt, 8, seed=11)
    path = tmp_path / "s.quads"
    with open(path, "wb") as f:
        np.array([qm.n_elem, qm.ng, qm.nf], np.int32).tofile(f)
        for a in (qm.vertices, qm.geom, qm.field):
            np.ascontiguousarray(a, np.float32).tofile(f)
    W, H = 96, 64
    aov, rays = tmp_path / "q.aov", tmp_path / "q.rays"
    r = subprocess.run([str(ROOT / "test-_amd" / "_build" / "femrt_render"), "--quads", str(path),
                        str(tmp_path / "q.ppm"), str(W), str(H), "--aov", str(aov), "--rays", str(rays)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"from {qm.n_elem} quads" in r.stdout
    a = np.fromfile(aov, AOV)
    ray = np.fromfile(rays, np.float32).reshape(6, W * H)
    og = orc.QuadOracle.from_mesh(qm).trace(ray)
    gh = dict(elem=a["elem"], face=a["face"], t=a["t"], rst=a["rst"].T.copy(), u=a["u"])
    rep = parity.compare(gh, None, og, None, [])
    print(json.dumps(rep), flush=True)
    assert rep["oracle_hits"] > 0.2 * W * H
    assert rep["ok"]
