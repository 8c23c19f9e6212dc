"""Curved triangles (the reference's TriElementData, fem/TriElement.cu; scene
setup raytracer.h:259-304).

The reference's triangle map is a sum over Modified_A(i, r)
Modified_B(i, j, s), j < my - i (EvaluateTriModal, fem/TriElement.cu:46-63):
collapsed coordinates, the edge s = +1 collapsed to a vertex.  So a triangle
is a quad map with one collapsed edge, and the library uploads that quad
(femrt_mesh_upload_tris) as a curved-quad surface.  The flat triangle is the
bilinear quad v0 v1 v2 v2 in the same coordinates.

* CPU: the oracle's triangle sums one mode at a time, the flat triangle's
  barycentric closed form, and the library's embedding
  (femrt_host_tri_to_quad, in double) against the oracle's direct sums.
* GPU: first hits of a triangulated terrain against the FP64 quad oracle on
  the uploaded embedding; vertical rays, which hit each cell's plane once;
  the field against the oracle's direct triangle sums."""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle.oracle as orc

TRI = np.array([[0.2, 0.1, 0.3], [1.4, 0.2, 0.1], [0.5, 1.3, 0.6]], np.float32)


def _fm():
    import importlib
    return importlib.import_module("test-_amd")


def test_tri_count_and_layout(oracle):
    fm = _fm()
    for m in [(2, 2), (3, 3), (4, 6), (6, 4)]:
        assert fm.tri_count(m) == orc.tri_count(m) == sum(max(m[1] - i, 0) for i in range(m[0]))
    m = (4, 5)
    r, s = 0.37, -0.61
    idx = 0
    for i in range(m[0]):
        for j in range(m[1] - i):
            g = np.zeros((1, 3, fm.tri_count(m)), np.float32)
            g[0, 2, idx] = 1.0
            x = orc.TriOracle(TRI[None], g, None, m).map(0, r, s)
            assert x[2] == pytest.approx(orc.modA(i, r) * orc.modB(i, j, s), abs=1e-14)
            assert x[0] == 0 and x[1] == 0
            idx += 1


def test_flat_triangle_barycentric(oracle):
    T = orc.TriOracle(TRI[None])
    rng = np.random.default_rng(2)
    for r, s in rng.uniform(-1, 1, (20, 2)):
        b1, b2 = (1 + r) * (1 - s) / 4, (1 + s) / 2
        assert b1 >= 0 and b2 >= 0 and b1 + b2 <= 1 + 1e-15   # the square covers the triangle
        np.testing.assert_allclose(T.map(0, r, s), (1 - b1 - b2) * TRI[0] + b1 * TRI[1] + b2 * TRI[2], atol=1e-14)
    np.testing.assert_allclose(T.map(0, -1, -1), TRI[0], atol=1e-7)
    np.testing.assert_allclose(T.map(0, 1, -1), TRI[1], atol=1e-7)
    for r in (-1.0, 0.0, 1.0):
        np.testing.assert_allclose(T.map(0, r, 1), TRI[2], atol=1e-7)


@pytest.mark.parametrize("gm,fm_", [((2, 2), (3, 3)), ((3, 3), (5, 5)), ((4, 5), (6, 3)), ((7, 7), (9, 9))])
def test_embedding_equals_direct_triangle_sums(oracle, gm, fm_):
    fm = _fm()
    rng = np.random.default_rng(sum(gm) * 10 + sum(fm_))
    n = 3
    v = rng.uniform(-1, 1, (n, 3, 3)).astype(np.float32)
    g = rng.uniform(-1, 1, (n, 3, fm.tri_count(gm))).astype(np.float32)
    f = rng.uniform(-1, 1, (n, fm.tri_count(fm_))).astype(np.float32)
    qv, qg, qf, ng, nf = fm.TriMesh(v, g, f, gm, fm_).to_quad()
    assert ng == max(gm) and nf == max(fm_)
    T = orc.TriOracle(v, g, f, gm, fm_)
    for e in range(n):
        for r, s in rng.uniform(-1, 1, (8, 2)):
            A = np.array([orc.modA(a, r) for a in range(max(ng, nf))])
            B = np.array([orc.modA(a, s) for a in range(max(ng, nf))])
            for c in range(3):
                xq = B[:ng] @ qg[e, c].reshape(ng, ng) @ A[:ng]       # [j][i]
                assert xq == pytest.approx(T.map(e, r, s)[c], abs=1e-12)
            uq = B[:nf] @ qf[e].reshape(nf, nf) @ A[:nf]
            assert uq == pytest.approx(T.field_at(e, r, s), abs=1e-12)


def test_flat_embedding_is_the_collapsed_quad(oracle):
    fm = _fm()
    qv, qg, qf, ng, nf = fm.TriMesh(TRI[None]).to_quad()
    assert ng == 2 and nf == 1
    np.testing.assert_array_equal(qv[0], TRI[[0, 1, 2, 2]])
    np.testing.assert_array_equal(qg[0].T, TRI[[0, 1, 2, 2]])


# ---------------------------------------------------------------- GPU


def _quad_oracle(tm):
    fm = _fm()
    qv, qg, qf, ng, nf = tm.to_quad()
    return orc.QuadOracle.from_mesh(fm.QuadMesh(qv, qg.astype(np.float32), qf.astype(np.float32), ng, nf))


@pytest.mark.gpu
@pytest.mark.parametrize("builder", ["host_sah", "gpu_lbvh"])
def test_tris_first_hit_parity(femrt, oracle, gpu_ctx, builder):
    import oracle.parity as parity
    tm = femrt.gen_tris(12, fm=(5, 5), amp=0.4, seed=3)
    mesh = femrt.Mesh.from_tris(gpu_ctx, tm, build_bvh=False)
    mesh.build_bvh(femrt.BVH_HOST_SAH if builder == "host_sah" else femrt.BVH_GPU_LBVH)
    W, H = 160, 120
    fr = femrt.Frame(gpu_ctx, W * H)
    fr.camera_rays(femrt.auto_camera([0, 0, -0.1], [1, 1, 0.2]), W, H)
    fr.trace(mesh, femrt.trace_params([], iso_enable=False))
    gh = fr.hits()
    rays = fr.get_rays()
    og = _quad_oracle(tm).trace(rays)
    rep = parity.compare(gh, None, og, None, [])
    print(json.dumps({"triangles": tm.n_elem, "builder": builder, "report": rep}), flush=True)
    assert rep["oracle_hits"] > 0.1 * W * H
    assert rep["ok"]
    h = np.flatnonzero(gh["elem"] >= 0)
    T = orc.TriOracle(tm.vertices, None, tm.field, (0, 0), tm.fm)
    sel = h[:: max(1, len(h) // 300)]
    du = max(abs(gh["u"][k] - T.field_at(gh["elem"][k], gh["rst"][0, k], gh["rst"][1, k])) for k in sel)
    dx = max(np.max(np.abs(T.map(gh["elem"][k], gh["rst"][0, k], gh["rst"][1, k]) -
                           (rays[:3, k] + gh["t"][k] * rays[3:, k]))) for k in sel)
    print(json.dumps({"field_vs_direct_max": float(du), "point_vs_direct_max": float(dx)}), flush=True)
    assert du < 1e-6 and dx < 1e-6
    fr.close()
    mesh.close()


@pytest.mark.gpu
def test_tris_vertical_rays_hit_each_cell_once(femrt, gpu_ctx):
    n = 8
    tm = femrt.gen_tris(n, fm=(3, 3), amp=0.4, seed=9)
    mesh = femrt.Mesh.from_tris(gpu_ctx, tm)
    rng = np.random.default_rng(4)
    k = 4096
    x, y = rng.uniform(0.002, 0.998, k), rng.uniform(0.002, 0.998, k)
    rays = np.stack([x, y, np.full(k, 3.0), np.zeros(k), np.zeros(k), -np.ones(k)]).astype(np.float32)
    fr = femrt.Frame(gpu_ctx, k)
    fr.set_rays(rays)
    fr.trace(mesh, femrt.trace_params([], iso_enable=False))
    gh = fr.hits()
    assert np.all(gh["elem"] >= 0)
    # the cell's triangle: lower-left when the in-cell offsets sum below one
    x32, y32 = rays[0].astype(np.float64) * n, rays[1].astype(np.float64) * n
    a, b = np.floor(y32).astype(int), np.floor(x32).astype(int)
    fx, fy = x32 - b, y32 - a
    clear = (np.abs(fx + fy - 1) > 1e-4) & (np.minimum(fx, fy) > 1e-4) & (np.maximum(fx, fy) < 1 - 1e-4)
    want = 2 * (a * n + b) + (fx + fy > 1)
    assert np.all(gh["elem"][clear] == want[clear])
    # the hit height is the plane of that triangle through its vertices
    V = tm.vertices.astype(np.float64)[gh["elem"]]
    nrm = np.cross(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0])
    px, py = rays[0].astype(np.float64), rays[1].astype(np.float64)
    z = V[:, 0, 2] - (nrm[:, 0] * (px - V[:, 0, 0]) + nrm[:, 1] * (py - V[:, 0, 1])) / nrm[:, 2]
    np.testing.assert_allclose(3.0 - gh["t"], z, atol=1e-12)
    with pytest.raises(RuntimeError, match="surface"):
        fr.trace(mesh, femrt.trace_params([0.1]))
    fr.close()
    mesh.close()


@pytest.mark.gpu
def test_cpp_scene_add_curved_tri_render(femrt, tmp_path):
    """Scene::add_curved_tri through the command-line renderer, against the
    quad oracle on the embedding, on the renderer's own rays."""
    import subprocess

    import oracle.parity as parity
    from conftest import ROOT
    from test_dropin import AOV
    tm = femrt.gen_tris(6, fm=(3, 3), amp=0.4, seed=13)
    path = tmp_path / "s.tris"
    with open(path, "wb") as f:
        np.array([tm.n_elem, 0, 0, *tm.fm], np.int32).tofile(f)
        for a in (tm.vertices, tm.field):
            np.ascontiguousarray(a, np.float32).tofile(f)
    W, H = 96, 64
    aov, rays = tmp_path / "t.aov", tmp_path / "t.rays"
    r = subprocess.run([str(ROOT / "test-_amd" / "_build" / "femrt_render"), "--tris", str(path),
                        str(tmp_path / "t.ppm"), str(W), str(H), "--aov", str(aov), "--rays", str(rays)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"from {tm.n_elem} triangles" in r.stdout
    a = np.fromfile(aov, AOV)
    ray = np.fromfile(rays, np.float32).reshape(6, W * H)
    og = _quad_oracle(tm).trace(ray)
    gh = dict(elem=a["elem"], face=a["face"], t=a["t"], rst=a["rst"].T.copy(), u=a["u"])
    rep = parity.compare(gh, None, og, None, [])
    print(json.dumps(rep), flush=True)
    assert rep["oracle_hits"] > 0.05 * W * H
    assert rep["ok"]
