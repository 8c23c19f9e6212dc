"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the
same seeded inputs (same FP32 coefficients, same FP32 rays)."""
from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _parity():
    import oracle.parity as parity
    return parity


def _mesh_c1(femrt, n=4):
    # config-1 style: structured 4^3, Q=3, sinusoidal face displacement 0.05,
    # field order 4
    return femrt.gen_structured_hex(n, 3, 4, pert=1, amp=0.05, seed=2024)


def test_eval_field_matches_oracle(femrt, oracle, gpu_ctx):
    hm = _mesh_c1(femrt)
    mesh = femrt.Mesh(gpu_ctx, hm)
    orc = oracle.Oracle.from_mesh(hm)
    rng = np.random.default_rng(10)
    n = 500
    elem = rng.integers(0, hm.n_elem, n).astype(np.int32)
    rst = rng.uniform(-1, 1, (3, n)).astype(np.float32)
    out64 = mesh.eval_field(elem, rst, precision=1)
    out32 = mesh.eval_field(elem, rst, precision=0)
    for i in range(n):
        u, _, gw = orc.field_at(elem[i], rst[:, i].astype(np.float64))
        assert out64[0, i] == pytest.approx(u, abs=1e-12)
        np.testing.assert_allclose(out64[1:, i], gw, rtol=1e-10, atol=1e-10)
        assert out32[0, i] == pytest.approx(u, abs=2e-6)
        np.testing.assert_allclose(out32[1:, i], gw, rtol=1e-4, atol=1e-4)
    mesh.close()


def test_point_inversion_matches_oracle(femrt, oracle, gpu_ctx):
    hm = _mesh_c1(femrt)
    mesh = femrt.Mesh(gpu_ctx, hm)
    orc = oracle.Oracle.from_mesh(hm)
    rng = np.random.default_rng(11)
    n = 400
    elem = rng.integers(0, hm.n_elem, n).astype(np.int32)
    x = np.zeros((3, n), np.float32)
    for i in range(n):
        x[:, i] = orc.map(elem[i], rng.uniform(-0.95, 0.95, 3))[0]
    rst, it = mesh.invert(elem, x)
    assert np.all(it > 0)
    for i in range(n):
        want, k = orc.invert(elem[i], x[:, i].astype(np.float64))
        assert k > 0
        np.testing.assert_allclose(rst[:, i], want, atol=1e-11)
    mesh.close()


def _trace_and_compare(femrt, oracle, ctx, hm, W, H, isos, guess, flags=1):
    mesh = femrt.Mesh(ctx, hm)
    lo = hm.vertices.reshape(-1, 3).min(0)
    hi = hm.vertices.reshape(-1, 3).max(0)
    cam = femrt.auto_camera(lo, hi)
    fr = femrt.Frame(ctx, W * H)
    fr.camera_rays(cam, W, H)
    fr.trace(mesh, femrt.trace_params(isos, guess=guess, flags=flags))
    gh = fr.hits()
    giso = fr.iso(len(isos))
    rays = fr.get_rays()
    orc = oracle.Oracle.from_mesh(hm)
    og, oiso = orc.trace(rays, isos)
    rep = _parity().compare(gh, giso, og, oiso, isos)
    fr.close()
    mesh.close()
    return rep


@pytest.mark.parametrize("guess,flags", [(0, 1), (1, 1), (0, 17), (1, 17)],
                         ids=["plain", "warm", "plain_coop", "warm_coop"])
def test_trace_config1_geometry_and_iso(femrt, oracle, gpu_ctx, guess, flags):
    hm = _mesh_c1(femrt)
    isos = [0.0]
    rep = _trace_and_compare(femrt, oracle, gpu_ctx, hm, 64, 48, isos, guess, flags)
    print(json.dumps(rep, indent=1))
    assert rep["oracle_hits"] > 200
    assert rep["ok"], rep


@pytest.mark.parametrize("flags", [1, 3, 17], ids=["persistent", "binned", "persistent_coop"])
def test_trace_three_isovalues_curved_q4(femrt, oracle, gpu_ctx, flags):
    """The walk schedules: one persistent kernel (sequential or cooperative
    scan), element-binned rounds."""
    hm = femrt.gen_structured_hex(4, 4, 6, pert=2, amp=0.08, seed=99)
    isos = [-0.1, 0.0, 0.1]
    rep = _trace_and_compare(femrt, oracle, gpu_ctx, hm, 48, 32, isos, 1, flags)
    print(json.dumps(rep, indent=1))
    assert rep["oracle_hits"] > 100
    assert rep["ok"], rep


def test_walk_across_rotated_face_orientations(femrt, oracle, gpu_ctx, tmp_path):
    """Elements whose local frames are rotated against their neighbours (as
    Nektar++ meshes have them): the walk maps the exit face coordinates into
    the neighbour's face frame.  3 x 1 x 2 trilinear hexes read by the XML
    loader, every other element's corner list rotated about its t axis."""
    nx, nz = 3, 2
    pid = lambda i, j, k: (k * 2 + j) * (nx + 1) + i  # noqa: E731
    verts = [(i / nx * 1.0, j * 0.5, k * 0.5) for k in range(nz + 1) for j in range(2) for i in range(nx + 1)]
    hexes = []
    for k in range(nz):
        for i in range(nx):
            h = [pid(i, 0, k), pid(i + 1, 0, k), pid(i + 1, 1, k), pid(i, 1, k),
                 pid(i, 0, k + 1), pid(i + 1, 0, k + 1), pid(i + 1, 1, k + 1), pid(i, 1, k + 1)]
            if (i + k) % 2:
                h = h[1:4] + h[0:1] + h[5:8] + h[4:5]   # rotate the local frame by 90 degrees
            hexes.append(h)
    vx = "\n".join(f'<V ID="{n}"> {x} {y} {z} </V>' for n, (x, y, z) in enumerate(verts))
    el = "\n".join(f'<H ID="{n}"> {" ".join(map(str, h))} </H>' for n, h in enumerate(hexes))
    xml = tmp_path / "rot.xml"
    xml.write_text(f'<NEKTAR><GEOMETRY DIM="3" SPACE="3"><VERTEX>{vx}</VERTEX><ELEMENT>{el}</ELEMENT>'
                   f'</GEOMETRY><EXPANSIONS><E COMPOSITE="C[0]" NUMMODES="4" FIELDS="u"/></EXPANSIONS></NEKTAR>')
    hm = femrt.load_xml_fld(xml, None)
    # every shared face seen from both sides
    assert (hm.adjacency[..., 0] >= 0).sum() == 2 * (nz * (nx - 1) + (nz - 1) * nx)
    # a smooth field shared across faces: the modal coefficients of
    # u = X(rst) . (0.7, -0.4, 1.1) per element (linear in the vertices)
    g = hm.geom.reshape(hm.n_elem, 3, hm.ng ** 3)
    lin = 0.7 * g[:, 0] - 0.4 * g[:, 1] + 1.1 * g[:, 2]
    f = np.zeros((hm.n_elem, 4, 4, 4), np.float32)
    f[:, :2, :2, :2] = lin.reshape(hm.n_elem, 2, 2, 2)
    hm.field = np.ascontiguousarray(f.reshape(hm.n_elem, -1))
    isos = [0.35, 0.6, 0.9]
    rep = _trace_and_compare(femrt, oracle, gpu_ctx, hm, 64, 48, isos, 1)
    print(json.dumps(rep, indent=1))
    assert rep["oracle_hits"] > 300
    assert all(r["gpu_found"] > 50 for r in rep["iso"])
    assert rep["ok"], rep


def test_scan_skip_by_field_bounds_keeps_parity(femrt, oracle, gpu_ctx):
    """Field u = z (the geometry's own z coefficients, so the isosurfaces are
    the planes z = 0.3 and z = 0.55 through a curved mesh): most elements'
    field bounds exclude the open isovalues and their scans are skipped; the
    results still match the oracle.  So do they when runs of such elements
    are passed by a BVH query (the default) and when every element is
    walked (FEMRT_TRACE_NO_SKIP_JUMP), which walks more elements."""
    hm = femrt.gen_structured_hex(4, 4, 4, pert=2, amp=0.08, seed=31)
    assert hm.ng == hm.nf
    hm.field = np.ascontiguousarray(hm.geom_blocks()[:, 2, :])
    isos = [0.3, 0.55]
    mesh = femrt.Mesh(gpu_ctx, hm)
    W, H = 64, 48
    fr = femrt.Frame(gpu_ctx, W * H)
    fr.camera_rays(femrt.auto_camera([0, 0, 0], [1, 1, 1]), W, H)
    fr.counters(True)
    fr.trace(mesh, femrt.trace_params(isos, guess=femrt.GUESS_WARM))
    cnt = fr.counters(False)
    og, oiso = oracle.Oracle.from_mesh(hm).trace(fr.get_rays(), isos)
    rep = _parity().compare(fr.hits(), fr.iso(len(isos)), og, oiso, isos)
    print(json.dumps(rep), cnt["scan_skipped"], cnt["elems_marched"])
    assert rep["ok"], rep
    assert rep["oracle_hits"] > 500 and all(r["oracle_found"] > 100 for r in rep["iso"])
    assert cnt["scan_skipped"] > 0.3 * cnt["elems_marched"]
    assert cnt["walk_jumps"] > 0
    fr.counters(True)
    fr.trace(mesh, femrt.trace_params(isos, guess=femrt.GUESS_WARM, flags=1 | femrt.TRACE_NO_SKIP_JUMP))
    cnt_w = fr.counters(False)
    rep_w = _parity().compare(fr.hits(), fr.iso(len(isos)), og, oiso, isos)
    print(json.dumps({"walk_every_element": rep_w}), cnt_w["elems_marched"], cnt_w["walk_jumps"])
    assert rep_w["ok"], rep_w
    assert cnt_w["walk_jumps"] == 0 and cnt_w["elems_marched"] > cnt["elems_marched"]
    fr.close()
    mesh.close()


def test_face_crossed_twice_gpu_takes_the_nearer_crossing(femrt, gpu_ctx):
    """Nearly horizontal rays below z = 0 cross the curved bottom face of the
    reference's Hex_CurvedFace fixture twice (into the bowl and out again).
    The GPU's hit must be the nearer crossing, which dense sampling of the
    closed-form face finds (tests/test_oracle.py), with warm guesses, both
    first-hit forms; the reference's nine seeds (plain) are reported."""
    from conftest import GOLDEN
    from test_oracle import _bowl_crossings
    hm = femrt.load_xml_fld(GOLDEN / "Hex_CurvedFace.xml", GOLDEN / "Hex_CurvedFace.fld")
    mesh = femrt.Mesh(gpu_ctx, hm)
    cases = [(-0.05, 0.5, 0.0, 0.0), (-0.12, 0.43, 0.07, 0.01), (-0.2, 0.6, -0.05, -0.02), (-0.27, 0.5, 0.0, 0.015)]
    o = np.array([[-0.5, y0, z0] for z0, y0, _, _ in cases], np.float32)
    d = np.array([[1.0, dy, dz] for _, _, dy, dz in cases], np.float32)
    rays = np.ascontiguousarray(np.concatenate([o.T, d.T]), np.float32)
    near = np.array([_bowl_crossings(o[i].astype(float), d[i].astype(float))[0] for i in range(len(cases))])
    fr = femrt.Frame(gpu_ctx, len(cases))
    fr.set_rays(rays)
    out = {}
    for name, guess, flags in (("warm", femrt.GUESS_WARM, 1), ("warm_nested", femrt.GUESS_WARM, 1 | femrt.TRACE_NESTED_FIRST_HIT),
                               ("plain", femrt.GUESS_PLAIN, 1)):
        fr.trace(mesh, femrt.trace_params([], guess=guess, iso_enable=False, flags=flags))
        h = fr.hits()
        out[name] = {"face": h["face"].tolist(), "t_err": np.abs(h["t"] - near).tolist()}
        if name != "plain":
            assert np.all(h["elem"] == 0) and np.all(h["face"] == 0), out
            np.testing.assert_allclose(h["t"], near, rtol=1e-9, err_msg=name)
    print(json.dumps(out), flush=True)
    fr.close()
    mesh.close()
