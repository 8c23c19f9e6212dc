"""Brute-force first crossings of u = iso along a ray, independent of the
oracle's per-element face roots and scan: dense samples along the ray, each
located by FP64 point inversion in the elements of the neighbouring cells of
a structured n^3 grid (the element whose preimage lies in the closed cube),
the first sign change of u - iso between located neighbours, bisection.

TEST INFRASTRUCTURE ONLY: the checker of the oracle's isosurface search.
"""
from __future__ import annotations

import numpy as np


class Locator:
    """Point location in a structured n^3 mesh of the unit cube."""

    def __init__(self, orc, hm, n):
        self.orc, self.n = orc, n
        cen = hm.vertices.reshape(-1, 8, 3).mean(axis=1)
        cell = np.clip(np.floor(cen * n).astype(int), 0, n - 1)
        self.grid = -np.ones((n, n, n), int)
        self.grid[cell[:, 0], cell[:, 1], cell[:, 2]] = np.arange(len(cen))
        assert (self.grid >= 0).all()

    def locate(self, x, hint=-1):
        cands = [hint] if hint >= 0 else []
        c = np.floor(np.asarray(x) * self.n).astype(int)
        for dz in (0, -1, 1):
            for dy in (0, -1, 1):
                for dx in (0, -1, 1):
                    i, j, k = c[0] + dx, c[1] + dy, c[2] + dz
                    if 0 <= i < self.n and 0 <= j < self.n and 0 <= k < self.n:
                        cands.append(int(self.grid[i, j, k]))
        for e in cands:
            rst, it = self.orc.invert(e, x)
            if it >= 0 and np.max(np.abs(rst)) <= 1 + 1e-9:
                return e, rst
        return -1, None

    def u_at(self, o, d, t, hint=-1):
        e, rst = self.locate(o + t * d, hint)
        return (e, self.orc.field_at(e, rst)[0]) if e >= 0 else (-1, np.nan)


def _u_in(loc, e, x):
    """u at x from element e's own preimage (None when x has none in e)."""
    rst, it = loc.orc.invert(e, x)
    if it < 0 or np.max(np.abs(rst)) > 1 + 1e-9:
        return None
    return loc.orc.field_at(e, rst)[0]


def first_crossings(loc, o, d, isos, t0, t1, dt):
    """[(t, elem) or None] per isovalue: the first sign change of u - iso
    along [t0, t1] within one element.  Samples every dt are located; where
    consecutive samples lie in different elements the face point between them
    is bisected and u is evaluated there from both sides, so each element's
    run of values ends and starts on its faces (a field may jump across a
    face; a crossing next to a face is still inside its element's run).  The
    first sign change within a run is bisected to 1e-13."""
    o = np.asarray(o, float)
    d = np.asarray(d, float)
    ts = np.arange(t0, t1, dt)
    es, us, hint = [], [], -1
    for t in ts:
        e, u = loc.u_at(o, d, t, hint)
        hint = e if e >= 0 else hint
        es.append(e)
        us.append(u)
    # runs: (element, [(t, u), ...]) in ray order
    runs = []
    for i, t in enumerate(ts):
        e = es[i]
        if e < 0:
            continue
        if i > 0 and es[i - 1] >= 0 and es[i - 1] != e:
            ep = es[i - 1]
            a, b = ts[i - 1], t
            for _ in range(60):
                m = 0.5 * (a + b)
                if _u_in(loc, ep, o + m * d) is not None:
                    a = m
                else:
                    b = m
            ua = _u_in(loc, ep, o + a * d)
            if ua is not None:
                runs[-1][1].append((a, ua))
            ub = _u_in(loc, e, o + b * d)
            runs.append((e, [(b, ub)] if ub is not None else []))
        elif not runs or runs[-1][0] != e or (i > 0 and es[i - 1] < 0):
            runs.append((e, []))
        runs[-1][1].append((t, us[i]))
    out = []
    for iso in isos:
        hit = None
        for e, seq in runs:
            for q in range(1, len(seq)):
                (ta, ua), (tb, ub) = seq[q - 1], seq[q]
                if (ua - iso) * (ub - iso) <= 0:
                    a, b, ga = ta, tb, ua - iso
                    while b - a > 1e-13 * (1 + abs(b)):
                        m = 0.5 * (a + b)
                        um = _u_in(loc, e, o + m * d)
                        if um is None:
                            break
                        if (um - iso) * ga <= 0:
                            b = m
                        else:
                            a, ga = m, um - iso
                    hit = (0.5 * (a + b), int(e))
                    break
            if hit:
                break
        out.append(hit)
    return out
