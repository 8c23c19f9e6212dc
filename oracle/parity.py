"""Parity comparison of GPU trace results against the FP64 oracle.

TEST INFRASTRUCTURE ONLY (used by tests/, __graft_entry__.smoke() and the
checks of bench.py).  Tolerances are the ones the acceptance bar states:
hit distance 1e-6 relative, reference coordinates 1e-7, field and isosurface
values 1e-6.  Tie pixels, where the element or face id is not determined to
the working precision, are classified by the oracle and excused from the id
check only (their distance is still checked when both sides hit).
"""
from __future__ import annotations

import numpy as np

T_REL = 1e-6
RST_ABS = 1e-7
VAL_ABS = 1e-6
# a ray is a tie pixel when another (element, face) candidate lies within this
# relative distance of the nearest hit, or the accepted/rejected face
# coordinate is within TIE_MARGIN of the acceptance bound
TIE_T = 1e-6
TIE_MARGIN = 1e-5
# an isovalue is tie-classified in a ray when some scan sample of g = u - iso
# before the oracle's first crossing (in any element the ray crossed; all of
# them when there is none) is within this of zero, or within 4e-6 t of its
# zero along the ray (|g| < 4e-6 t |dg/dt|: the FP32 scan places its samples
# to a few 1e-7 t): the sign decisions there are not determined.  The oracle
# reports both tests on one scale (minabs < TIE_ISO).
TIE_ISO = 1e-5


def geometry_ties(og):
    """Hit rays: another candidate within TIE_T of the hit distance, or the hit
    within TIE_MARGIN of its face boundary.  Miss rays: a converged root just
    outside a face (within TIE_MARGIN of acceptance)."""
    t = og["t"]
    hit = og["elem"] >= 0
    near_hit = (np.abs(og["t2"] - t) <= TIE_T * np.abs(t)) | (og["margin"] < TIE_MARGIN)
    return near_hit & hit | (og["nearmiss"] < TIE_MARGIN)


def compare(gh, giso, og, oiso, isos):
    """Returns a dict of counts and worst errors; `ok` tells whether the bar
    is met."""
    rep = {}
    tie = geometry_ties(og)
    ho, hg = og["elem"] >= 0, gh["elem"] >= 0
    rep["rays"] = int(ho.size)
    rep["oracle_hits"] = int(ho.sum())
    rep["gpu_hits"] = int(hg.sum())
    rep["ties"] = int(tie.sum())
    rep["hit_miss_mismatch"] = int(((ho != hg) & ~tie).sum())
    both = ho & hg
    trel = np.abs(gh["t"] - og["t"]) / np.maximum(np.abs(og["t"]), 1e-300)
    nt = both & ~tie
    rep["t_rel_max"] = float(trel[nt].max()) if nt.any() else 0.0
    # at tie pixels two crossings lie within TIE_T of each other, so either
    # may be reported; their distance agreement is shown, not gated
    rep["t_rel_max_ties"] = float(trel[both & tie].max()) if (both & tie).any() else 0.0
    same = both & (gh["elem"] == og["elem"]) & (gh["face"] == og["face"])
    rep["id_mismatch"] = int((both & ~same & ~tie).sum())
    rep["id_mismatch_ties"] = int((both & ~same & tie).sum())
    if same.any():
        drst = np.abs(gh["rst"][:, same] - og["rst"][:, same]).max()
        du = np.abs(gh["u"][same] - og["u"][same]).max()
    else:
        drst = du = 0.0
    rep["rst_abs_max"] = float(drst)
    rep["u_abs_max"] = float(du)
    ok = (rep["hit_miss_mismatch"] == 0 and rep["id_mismatch"] == 0 and
          rep["t_rel_max"] <= T_REL and rep["rst_abs_max"] <= RST_ABS and
          rep["u_abs_max"] <= VAL_ABS)
    iso_rep = []
    for k, iso in enumerate(isos):
        fo, fg = oiso["elem"][k] >= 0, giso["elem"][k] >= 0
        # the walk starts at the geometry hit: a tie there ties the iso result
        itie = (oiso["minabs"][k] < TIE_ISO) | tie
        r = {"iso": float(iso), "oracle_found": int(fo.sum()), "gpu_found": int(fg.sum()),
             "found_mismatch": int(((fo != fg) & ~itie).sum()), "ties": int(itie.sum())}
        b = fo & fg
        if b.any():
            tr = np.abs(giso["t"][k][b] - oiso["t"][k][b]) / np.abs(oiso["t"][k][b])
            r["t_rel_max"] = float(tr.max())
            r["t_rel_bad"] = int((tr > T_REL).sum())
            r["t_rel_bad_nontie"] = int(((tr > T_REL) & ~itie[b]).sum())
            # u of rays without a crossing is NaN (not reported): compare found rays only
            du = np.zeros(fg.shape)
            du[fg] = np.abs(giso["u"][k][fg] - iso)
            r["u_minus_iso_max"] = float(du[fg & ~itie].max()) if (fg & ~itie).any() else 0.0
            # near-tangent crossings (tie class) are reported, not gated
            r["u_minus_iso_max_ties"] = float(du[fg & itie].max()) if (fg & itie).any() else 0.0
            se = b & (giso["elem"][k] == oiso["elem"][k])
            r["elem_mismatch"] = int((b & ~se & ~itie).sum())
            dr = np.abs(giso["rst"][k] - oiso["rst"][k]).max(axis=0)
            sn = se & ~itie
            r["rst_abs_max"] = float(dr[sn].max()) if sn.any() else 0.0
            r["rst_abs_max_ties"] = float(dr[se & itie].max()) if (se & itie).any() else 0.0
        else:
            r.update(t_rel_max=0.0, t_rel_bad=0, t_rel_bad_nontie=0, u_minus_iso_max=0.0,
                     elem_mismatch=0, rst_abs_max=0.0, rst_abs_max_ties=0.0)
        r["ok"] = (r["found_mismatch"] == 0 and r["t_rel_bad_nontie"] == 0 and
                   r["u_minus_iso_max"] <= VAL_ABS and r["elem_mismatch"] == 0 and
                   r["rst_abs_max"] <= RST_ABS)
        ok = ok and r["ok"]
        iso_rep.append(r)
    rep["iso"] = iso_rep
    rep["ok"] = bool(ok)
    return rep
