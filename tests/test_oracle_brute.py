"""The oracle's isosurface search pinned by brute force along the ray
(tests/brute_force.py): on a small curved mesh with a multi-frequency field,
every isovalue's first crossing the oracle reports must be the one dense
point location and bisection find, and it must report none where they find
none.  The rays include the configurations that defeat per-element face
roots: rays through element edges (in the diagonal plane x = y, which holds
the z-edges of the grid) and rays inside a grid plane (x = 1/2, where the
curved faces of two element columns meet).  Crossings where a sample of u -
iso lies within the oracle's tie margin of zero are excused (the scan's
sign decision there is a tie)."""
from __future__ import annotations

import numpy as np
import pytest

from brute_force import Locator, first_crossings

N = 6
ISOS = [-0.5, -0.15, 0.15, 0.5]


def _rays():
    rng = np.random.default_rng(77)
    rays = []
    for _ in range(6):   # diagonal plane x = y: through the z-edges
        a = rng.uniform(-0.3, -0.1)
        dz = rng.uniform(-0.4, 0.4)
        rays.append([a, a, rng.uniform(0.2, 0.8), 1.0, 1.0, dz])
    for _ in range(5):   # inside the grid plane x = 1/2
        rays.append([0.5, rng.uniform(-0.4, -0.1), rng.uniform(0.1, 0.9), 0.0, 1.0, rng.uniform(-0.3, 0.3)])
    for _ in range(6):   # generic directions
        o = rng.uniform(-0.5, 1.5, 3)
        o[0] = -0.4
        tgt = rng.uniform(0.2, 0.8, 3)
        rays.append([*o, *(tgt - o)])
    r = np.array(rays, np.float64)
    r[:, 3:] /= np.linalg.norm(r[:, 3:], axis=1, keepdims=True)
    return np.ascontiguousarray(r.T.astype(np.float32))


def test_oracle_iso_search_matches_brute_force_on_edge_and_face_rays(femrt, oracle):
    import oracle.parity as parity
    base = femrt.gen_structured_hex(N, 4, 2, pert=3, amp=0.15, seed=31)
    hm = femrt.gen_field_waves(base, 4, n_waves=8, seed=32)
    orc = oracle.Oracle.from_mesh(hm)
    rays = _rays()
    og, oi = orc.trace(rays, ISOS)
    loc = Locator(orc, hm, N)
    checked = found = 0
    for q in range(rays.shape[1]):
        o, d = rays[:3, q].astype(np.float64), rays[3:, q].astype(np.float64)
        bf = first_crossings(loc, o, d, ISOS, 0.0, 3.0, 1e-3)
        for k in range(len(ISOS)):
            tie = oi["minabs"][k][q] < parity.TIE_ISO
            if tie:
                continue
            checked += 1
            if bf[k] is None:
                assert oi["elem"][k][q] < 0, (q, ISOS[k], oi["t"][k][q])
                continue
            found += 1
            assert oi["elem"][k][q] >= 0, (q, ISOS[k], bf[k])
            assert oi["t"][k][q] == pytest.approx(bf[k][0], rel=1e-9), (q, ISOS[k])
    assert checked >= 60 and found >= 30
