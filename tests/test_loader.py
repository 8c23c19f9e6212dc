"""Nektar++ XML/FLD reader and writer (host code, no GPU): the reference's
Hex_CurvedFace fixture against its closed forms, round trips of generated
meshes, the 8-vertex element form, FLD ID placement across several ELEMENTS
blocks, the two reference quirks behind flags, and error reporting."""
from __future__ import annotations

import base64
import itertools
import zlib

import numpy as np
import pytest

from conftest import ROOT, decode_fld

GOLD = ROOT / "tests" / "golden"


def _map_err(orc, e, fn, n=200, seed=0):
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(n):
        rst = rng.uniform(-1, 1, 3)
        x, _ = orc.map(e, rst)
        worst = max(worst, float(np.max(np.abs(x - fn(*rst)))))
    return worst


def _hcf_exact(r, s, t):
    return np.array([(r + 1) / 2, (s + 1) / 2, (t + 1) / 2 - 0.3 * (1 - r * r) * (1 - s * s) * (1 - t) / 2])


@pytest.mark.parametrize("flags,ng", [(0, 3), (2, 5)])
def test_hex_curved_face_fixture(femrt, oracle, flags, ng):
    hm = femrt.load_xml_fld(GOLD / "Hex_CurvedFace.xml", GOLD / "Hex_CurvedFace.fld", flags)
    assert (hm.n_elem, hm.ng, hm.nf) == (1, ng, 5)
    corners = np.array(list(itertools.product((0, 1), repeat=3)))[:, ::-1]  # z slowest
    want = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    np.testing.assert_array_equal(hm.vertices[0], want)
    del corners
    # the geometry is the closed form (degree 2 face, linear blend) to FP32
    orc = oracle.Oracle.from_mesh(hm)
    assert _map_err(orc, 0, _hcf_exact) < 3e-7
    # field = the 125 decoded doubles, rounded to FP32
    np.testing.assert_array_equal(hm.field[0], decode_fld(GOLD / "Hex_CurvedFace.fld").astype(np.float32))
    u, _, _ = orc.field_at(0, np.array([0.0, 0.0, 0.0]))
    assert u == pytest.approx(0.261498792600390, abs=1e-6)
    assert np.all(hm.adjacency == -1)


def test_vertex0_quirk_flag(femrt, tmp_path):
    text = (GOLD / "Hex_CurvedFace.xml").read_text()
    # shift every vertex and curved point by (1, 2, 3)
    import re

    def shift(m):
        v = [float(x) for x in m.group(2).split()]
        pts = [f"{v[i] + 1} {v[i + 1] + 2} {v[i + 2] + 3}" for i in range(0, len(v), 3)]
        return m.group(1) + " " + "  ".join(pts) + " " + m.group(3)
    a, b = text.index("<VERTEX>"), text.index("</VERTEX>")
    text = text[:a] + re.sub(r"(<V ID=\"\d+\">)([^<]*)(</V>)", shift, text[a:b]) + text[b:]
    text = re.sub(r"(<[EF] ID=\"\d+\" (?:EDGE|FACE)ID=[^>]*>)([^<]*)(</[EF]>)", shift, text)
    xml = tmp_path / "shifted.xml"
    xml.write_text(text)
    fld = GOLD / "Hex_CurvedFace.fld"
    hm = femrt.load_xml_fld(xml, fld)
    np.testing.assert_allclose(hm.vertices[0, 0], [1, 2, 3])
    hq = femrt.load_xml_fld(xml, fld, femrt.LOAD_VERTEX0_DEFECT)
    np.testing.assert_array_equal(hq.vertices[0, 0], [0, 0, 0])
    np.testing.assert_allclose(hq.vertices[0, 1:], hm.vertices[0, 1:])


@pytest.mark.parametrize("n,q,p,pert,amp", [(3, 4, 3, 2, 0.08), (2, 3, 4, 1, 0.05)])
def test_round_trip_generated_mesh(femrt, tmp_path, n, q, p, pert, amp):
    hm = femrt.gen_structured_hex(n, q, p, pert=pert, amp=amp, seed=17)
    xml, fld = tmp_path / "m.xml", tmp_path / "m.fld"
    femrt.save_xml_fld(hm, xml, fld)
    assert 'COMPRESSED="B64Z-LittleEndian"' in fld.read_text()
    back = femrt.load_xml_fld(xml, fld)
    assert (back.n_elem, back.ng, back.nf) == (hm.n_elem, hm.ng, hm.nf)
    np.testing.assert_array_equal(back.vertices, hm.vertices)
    np.testing.assert_array_equal(back.field, hm.field)
    np.testing.assert_array_equal(back.adjacency, hm.adjacency)
    # geometry rebuilt from the written face points: trilinear + face blends
    # reproduce the generator's construction up to FP32 rounding
    np.testing.assert_allclose(back.geom, hm.geom, atol=2e-6)


def _write_plain_mesh(path, verts, hexes_8, nummodes):
    vx = "\n".join(f'<V ID="{i}"> {x} {y} {z} </V>' for i, (x, y, z) in enumerate(verts))
    el = "\n".join(f'<H ID="{i}"> {" ".join(map(str, h))} </H>' for i, h in enumerate(hexes_8))
    path.write_text(f"""<?xml version="1.0"?>
<NEKTAR><GEOMETRY DIM="3" SPACE="3">
<!-- <VERTEX><V ID="0"> 9 9 9 </V></VERTEX> commented out -->
<VERTEX>{vx}</VERTEX>
<ELEMENT>{el}</ELEMENT>
<COMPOSITE><C ID="0"> H[0-{len(hexes_8) - 1}] </C></COMPOSITE>
</GEOMETRY>
<EXPANSIONS><E COMPOSITE="C[0]" NUMMODES="{nummodes}" FIELDS="u" TYPE="MODIFIED"/></EXPANSIONS>
</NEKTAR>""")


def test_vertex_form_ascii_and_id_placement(femrt, oracle, tmp_path):
    # two stacked trilinear (sheared) hexes given by 8 vertex ids each
    verts = [(0, 0, 0), (1, 0, 0), (1.2, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1.2, 1, 1.1), (0, 1, 1),
             (0, 0, 2), (1, 0, 2), (1.2, 1, 2.1), (0, 1, 2)]
    hexes = [(0, 1, 2, 3, 4, 5, 6, 7), (4, 5, 6, 7, 8, 9, 10, 11)]
    xml = tmp_path / "v.xml"
    _write_plain_mesh(xml, verts, hexes, 2)
    # FLD: element 1 first (ASCII), then element 0 (B64Z) in a second block
    c0 = np.arange(8, dtype=np.float64) + 0.5
    c1 = -np.arange(8, dtype=np.float64)
    z = base64.b64encode(zlib.compress(c0.astype("<f8").tobytes())).decode()
    fld = tmp_path / "v.fld"
    fld.write_text(f"""<NEKTAR>
<ELEMENTS FIELDS="u" SHAPE="Hexahedron" NUMMODESPERDIR="UNIORDER:2,2,2" ID="1">{' '.join(map(str, c1))}</ELEMENTS>
<ELEMENTS FIELDS="u" SHAPE="Hexahedron" NUMMODESPERDIR="UNIORDER:2,2,2" ID="0" COMPRESSED="B64Z-LittleEndian" BITSIZE="64">{z}</ELEMENTS>
</NEKTAR>""")
    hm = femrt.load_xml_fld(xml, fld)
    assert (hm.n_elem, hm.ng, hm.nf) == (2, 2, 2)
    np.testing.assert_array_equal(hm.field[0], c0.astype(np.float32))
    np.testing.assert_array_equal(hm.field[1], c1.astype(np.float32))
    assert hm.adjacency[0, 5].tolist() == [1, 0] and hm.adjacency[1, 0].tolist() == [0, 5]
    orc = oracle.Oracle.from_mesh(hm)
    V = np.array(verts, float)
    for e, h in enumerate(hexes):
        C = V[list(h)]

        def tri(r, s, t, C=C):
            w = [(1 - r) * (1 - s) * (1 - t), (1 + r) * (1 - s) * (1 - t), (1 + r) * (1 + s) * (1 - t),
                 (1 - r) * (1 + s) * (1 - t), (1 - r) * (1 - s) * (1 + t), (1 + r) * (1 - s) * (1 + t),
                 (1 + r) * (1 + s) * (1 + t), (1 - r) * (1 + s) * (1 + t)]
            return np.dot(np.array(w) / 8, C)
        assert _map_err(orc, e, tri, n=50) < 1e-6
    # no FLD: modes from the expansion, zero field
    hz = femrt.load_xml_fld(xml, None)
    assert hz.nf == 2 and not hz.field.any()


def test_errors_are_reported(femrt, tmp_path):
    with pytest.raises(femrt.FemrtError, match="EIO.*cannot read"):
        femrt.load_xml_fld(tmp_path / "missing.xml", None)
    bad = tmp_path / "bad.xml"
    bad.write_text("<NEKTAR><GEOMETRY></GEOMETRY></NEKTAR>")
    with pytest.raises(femrt.FemrtError, match="no vertices"):
        femrt.load_xml_fld(bad, None)
    fld = tmp_path / "bad.fld"
    fld.write_text('<ELEMENTS FIELDS="u" NUMMODESPERDIR="UNIORDER:2,2,2" ID="0" COMPRESSED="B64Z-LittleEndian">!!notzlib</ELEMENTS>')
    xml = tmp_path / "ok.xml"
    _write_plain_mesh(xml, [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
                      [(0, 1, 2, 3, 4, 5, 6, 7)], 2)
    with pytest.raises(femrt.FemrtError, match="base64"):
        femrt.load_xml_fld(xml, fld)


@pytest.mark.parametrize("reverse", [False, True])
def test_curved_edge_is_resampled_onto_the_edge(femrt, oracle, tmp_path, reverse):
    """One unit-cube hex whose edge from vertex 0 (0,0,0) to vertex 1 (1,0,0)
    is curved: 4 evenly spaced points on z = -0.2 s (1 - s) (the edge listed
    in either direction).  The loaded map along that edge (s = t = -1) is the
    parabola (degree 2, the cubic geometry order holds it exactly), every
    other GLL node keeps its trilinear position, and the other edges stay
    straight."""
    verts = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    edges = [(0, 1), (1, 2), (3, 2), (0, 3), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (7, 6), (4, 7)]
    if reverse:
        edges[0] = (1, 0)
    s = np.linspace(0, 1, 4)
    pts = np.stack([s, 0 * s, -0.2 * s * (1 - s)], 1)
    if reverse:
        pts = pts[::-1]
    vx = "\n".join(f'<V ID="{i}"> {x} {y} {z} </V>' for i, (x, y, z) in enumerate(verts))
    ex = "\n".join(f'<E ID="{i}"> {a} {b} </E>' for i, (a, b) in enumerate(edges))
    cx = " ".join(f"{v:.17g}" for v in pts.ravel())
    xml = tmp_path / "ce.xml"
    xml.write_text(f"""<?xml version="1.0"?>
<NEKTAR><GEOMETRY DIM="3" SPACE="3">
<VERTEX>{vx}</VERTEX>
<EDGE>{ex}</EDGE>
<ELEMENT><H ID="0"> 0 1 2 3 4 5 6 7 </H></ELEMENT>
<CURVED><E ID="0" EDGEID="0" NUMPOINTS="4" TYPE="PolyEvenlySpaced"> {cx} </E></CURVED>
<COMPOSITE><C ID="0"> H[0] </C></COMPOSITE>
</GEOMETRY>
<EXPANSIONS><E COMPOSITE="C[0]" NUMMODES="2" FIELDS="u" TYPE="MODIFIED"/></EXPANSIONS>
</NEKTAR>""")
    hm = femrt.load_xml_fld(xml, None)
    assert hm.ng == 4
    orc = oracle.Oracle.from_mesh(hm)
    for r in np.linspace(-1, 1, 21):
        x, _ = orc.map(0, [r, -1, -1])
        sp = (r + 1) / 2
        np.testing.assert_allclose(x, [sp, 0, -0.2 * sp * (1 - sp)], atol=2e-6)
    # the opposite edges (s = +1 or t = +1) stay straight
    for r in np.linspace(-1, 1, 11):
        for (sv, tv) in ((1, -1), (-1, 1), (1, 1)):
            x, _ = orc.map(0, [r, sv, tv])
            np.testing.assert_allclose(x, [(r + 1) / 2, (sv + 1) / 2, (tv + 1) / 2], atol=2e-6)
    # GLL nodes off the curved edge keep their trilinear positions
    g = np.array([-1.0, -1 / np.sqrt(5), 1 / np.sqrt(5), 1.0])
    for r, s_, t in itertools.product(g, g, g):
        if s_ == -1 and t == -1:
            continue
        x, _ = orc.map(0, [r, s_, t])
        np.testing.assert_allclose(x, [(r + 1) / 2, (s_ + 1) / 2, (t + 1) / 2], atol=2e-6)


def test_round_trip_at_32_cubed_is_exact_and_timed(femrt, tmp_path):
    """The 32^3 benchmark mesh (32768 hexes, Q=4, P=6: 86 MB of XML and 73 MB
    of FLD) written and read back: identical field coefficients, geometry to
    the modal fit's rounding, and both within a time bound (the reader maps
    the files and scans them once; the per-element fits run on all cores)."""
    import time
    hm = femrt.gen_structured_hex(32, 4, 6, pert=2, amp=0.08, seed=20240611)
    xml, fld = tmp_path / "m.xml", tmp_path / "m.fld"
    t0 = time.perf_counter()
    femrt.save_xml_fld(hm, xml, fld)
    t1 = time.perf_counter()
    h2 = femrt.load_xml_fld(xml, fld)
    t2 = time.perf_counter()
    print({"save_s": t1 - t0, "load_s": t2 - t1, "xml_mb": xml.stat().st_size / 1e6,
           "fld_mb": fld.stat().st_size / 1e6})
    assert (h2.n_elem, h2.ng, h2.nf) == (hm.n_elem, hm.ng, hm.nf)
    np.testing.assert_array_equal(h2.field, hm.field.reshape(h2.field.shape))
    assert np.abs(h2.geom - hm.geom.reshape(h2.geom.shape)).max() < 2e-6
    np.testing.assert_array_equal(h2.adjacency[..., 0] >= 0, hm.adjacency.reshape(h2.adjacency.shape)[..., 0] >= 0)
    assert t2 - t1 < 30 and t1 - t0 < 60
