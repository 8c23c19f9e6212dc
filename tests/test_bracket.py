"""Learned isosurface bracketing: the bracket network (32 -> 64 -> 64 -> 2
quantiles, BF16 MFMA) against its FP64 oracle, trained by the library's own
kernels on pairs the tracer produces, and the FP64 bisection endgame it
starts, against plain bisection of the scan's bracket and the default
Newton endgame (fem/RayIsosurface.cu:121-171 is the reference's endgame).

The workload is config 5's family at a size the oracle checks quickly: a
curved 12^3 hex mesh, Q=4, P=8 multi-frequency field, four isovalues; the
network is trained on another seed of the same family."""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle.mlp_oracle as mo

pytestmark = pytest.mark.gpu

ISOS = [-0.6, -0.2, 0.2, 0.6]


@pytest.mark.parametrize("n", [64, 100, 257])
def test_bracket_forward_and_gradient_match_oracle(femrt, gpu_ctx, n):
    spec = mo.BRACKET
    mlp = femrt.Mlp(gpu_ctx, seed=3, kind=femrt.MLP_BRACKET)
    assert mlp.n_param == spec.n_param == 6402
    p = mo.glorot_params(21, spec=spec).astype(np.float32)
    mlp.set_params(p)
    rng = np.random.default_rng(n)
    feats = rng.uniform(-1, 1, (n, spec.n_in)).astype(np.float32)
    y_gpu = mlp.forward(feats)
    y_orc, _ = mo.forward(p, feats, True, spec)
    assert np.max(np.abs(y_gpu - y_orc)) < 3e-3
    # targets away from the outputs, so no pinball kink separates GPU and oracle
    t = y_orc[:, 0] + rng.choice([-1, 1], n) * rng.uniform(0.05, 1.0, n)
    tgt = np.stack([t, t], 1).astype(np.float32)
    loss, g = mlp.loss_grad(feats, tgt)
    l_o, g_o = mo.loss_grad(p, feats, tgt, True, spec)
    assert loss == pytest.approx(l_o, rel=1e-3)
    assert np.linalg.norm(g - g_o) / np.linalg.norm(g_o) < 1e-2
    o = spec.offsets
    for name, a, b in (("W1", o["W1"], o["b1"]), ("b1", o["b1"], o["W2"]), ("W3", o["W3"], o["b3"]),
                       ("b3", o["b3"], o["end"])):
        assert np.linalg.norm(g[a:b] - g_o[a:b]) / np.linalg.norm(g_o[a:b]) < 2e-2, name
    mlp.close()


def _mesh(femrt, seed):
    hm = femrt.gen_structured_hex(12, 4, 2, pert=3, amp=0.15, seed=seed)
    return femrt.gen_field_waves(hm, 8, n_waves=8, seed=seed + 1)


def _frame(femrt, ctx, W, H, spp, seed):
    fr = femrt.Frame(ctx, W * H * spp)
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 1])
    cam12 = femrt.camera_basis(cam[:3], cam[3:6], cam[6:9], float(cam[9]), W / H)
    fr.camera_rays_jitter(cam12, W, H, spp, seed)
    return fr


def _trace(femrt, fr, mesh, flags):
    fr.counters(True)
    fr.trace(mesh, femrt.trace_params(ISOS, guess=femrt.GUESS_WARM, flags=1 | flags))
    cnt = fr.counters(False)
    return fr.hits(), fr.iso(len(ISOS)), cnt


def test_learned_bracket_endgame_parity_and_gain(femrt, oracle, gpu_ctx):
    import oracle.parity as parity
    ctx = gpu_ctx
    # training pairs from the tracer on four other meshes of the family (one
    # mesh's pairs share its few fields: a network fitted to them does not
    # carry over to another field)
    parts = []
    for sd in (901, 921, 941, 961):
        tmesh = femrt.Mesh(ctx, _mesh(femrt, sd))
        tfr = _frame(femrt, ctx, 400, 300, 2, sd + 1)
        tfr.trace(tmesh, femrt.trace_params(ISOS, guess=femrt.GUESS_WARM, flags=1 | femrt.TRACE_SEG_INFO))
        parts.append(tfr.bracket_pairs(tmesh, ISOS, 1 << 20))
        tfr.close()
        tmesh.close()
    feats = np.concatenate([p[0] for p in parts])
    tgt = np.concatenate([p[1] for p in parts])
    assert feats.shape[0] > 20000
    assert np.all((tgt[:, 0] >= -1e-6) & (tgt[:, 0] <= 1 + 1e-6))
    # inputs: rst at the bracket's ends, the secant weight in [0, 1], and the
    # low-order coefficients of u - iso scaled to unit maximum
    assert np.all(np.abs(feats[:, :6]) <= 1.01)
    assert np.all((feats[:, 6] >= 0) & (feats[:, 6] <= 1))
    assert np.all(np.abs(feats[:, 7:]).max(axis=1) == pytest.approx(1.0))
    mlp = femrt.Mlp(ctx, seed=5, kind=femrt.MLP_BRACKET)
    hold = len(parts[-1][0])    # the last mesh's pairs are held out
    losses = mlp.train_pairs(feats[:-hold], tgt[:-hold], batch=4096, epochs=4, lr=3e-3, lr_final=1e-4, seed=7)
    assert losses[-1] < 0.7 * losses[0]
    q = mlp.forward(feats[-hold:])
    lo, hi = np.minimum(q[:, 0], q[:, 1]), np.maximum(q[:, 0], q[:, 1])
    cover = float(np.mean((tgt[-hold:, 0] >= lo) & (tgt[-hold:, 0] <= hi)))
    width = float(np.median(hi - lo))
    assert 0.6 < cover < 0.98
    assert width < 0.5      # narrower than the scan's bracket

    hm = _mesh(femrt, 911)
    mesh = femrt.Mesh(ctx, hm)
    mesh.set_bracket_mlp(mlp)
    fr = _frame(femrt, ctx, 320, 240, 2, 912)
    # hit polish alone (no isovalues): its FP64 work is in every count below
    fr.counters(True)
    fr.trace(mesh, femrt.trace_params([], guess=femrt.GUESS_WARM, iso_enable=False))
    h_c = fr.counters(False)
    a_h, a_i, a_c = _trace(femrt, fr, mesh, 0)
    b_h, b_i, b_c = _trace(femrt, fr, mesh, femrt.TRACE_ISO_BISECT)
    c_h, c_i, c_c = _trace(femrt, fr, mesh, femrt.TRACE_ISO_LEARNED_BRACKET)
    found = int((a_i["elem"] >= 0).sum())
    assert found > 5000 and a_c["brackets"] >= found
    # the same crossings: the FP64 roots of the three endgames agree to the
    # bisection tolerance wherever they report one; a crossing may differ only
    # where the FP32 scan's sign decision is a tie (checked by the oracle)
    diff, odd = {}, np.zeros(a_i["elem"].shape[1], bool)
    for name, gi in (("bisect", b_i), ("learned", c_i)):
        de = (gi["elem"] != a_i["elem"]).any(axis=0)
        odd |= de
        f = (gi["elem"] >= 0) & (a_i["elem"] >= 0)
        trel = np.abs(gi["t"][f] - a_i["t"][f]) / np.abs(a_i["t"][f])
        diff[name] = {"rays_differ": int(de.sum()), "t_rel_max": float(trel.max()), "u_minus_iso_max": float(max(
            np.abs(gi["u"][k][gi["elem"][k] >= 0] - ISOS[k]).max() for k in range(len(ISOS))))}
        assert diff[name]["rays_differ"] <= 1e-3 * found, (name, diff[name])
        assert diff[name]["t_rel_max"] < 1e-9 and diff[name]["u_minus_iso_max"] <= 1e-7, (name, diff[name])
    # sampled rays of the learned endgame, and every ray where an endgame
    # differs, against the oracle at the bar
    rng = np.random.default_rng(3)
    crossing = np.flatnonzero((c_i["elem"] >= 0).any(axis=0))
    sel = np.union1d(rng.choice(crossing, 120, replace=False), np.flatnonzero(odd))
    rays = fr.get_rays()
    og, oi = oracle.Oracle.from_mesh(hm).trace(np.ascontiguousarray(rays[:, sel]), ISOS)
    sub = lambda d: {k: (v[..., sel] if isinstance(v, np.ndarray) else v) for k, v in d.items()}  # noqa: E731
    for name, (gh, gi) in (("learned", (c_h, c_i)), ("bisect", (b_h, b_i)), ("newton", (a_h, a_i))):
        rep = parity.compare(sub(gh), sub(gi), og, oi, ISOS)
        if not rep["ok"]:
            print(json.dumps({"endgame": name, "report": rep}), flush=True)
        assert rep["ok"], name
    # the network on the evaluation mesh's own brackets (the learned trace
    # recorded their inputs): coverage and width of the predicted interval
    ef, et = fr.bracket_pairs(mesh, ISOS, 1 << 20)
    q = mlp.forward(ef)
    lo, hi = np.minimum(q[:, 0], q[:, 1]), np.maximum(q[:, 0], q[:, 1])
    eval_cover = float(np.mean((et[:, 0] >= lo) & (et[:, 0] <= hi)))
    eval_width = float(np.median(np.clip(hi, 0, 1) - np.clip(lo, 0, 1)))
    # what the network gains: FP64 field evaluations and inversion iterations
    # of the endgame per bracket (the hit polish's share taken out), learned
    # bracket against plain bisection
    nb = a_c["brackets"]

    def per(c, k):
        return (c[k] - h_c[k]) / nb
    out = {"brackets": nb, "holdout_cover": cover, "holdout_median_width": width,
           "eval_mesh_cover": eval_cover, "eval_mesh_median_width": eval_width,
           "field64_per_bracket": {"newton": per(a_c, "field64_evals"), "bisect": per(b_c, "field64_evals"),
                                   "learned_bisect": per(c_c, "field64_evals")},
           "geo64_per_bracket": {"newton": per(a_c, "geo64_it"), "bisect": per(b_c, "geo64_it"),
                                 "learned_bisect": per(c_c, "geo64_it")},
           "agreement": diff}
    print(json.dumps({"learned_bracket": out}), flush=True)
    assert per(c_c, "field64_evals") < per(b_c, "field64_evals") - 1.0
    for o in (fr, mesh, mlp):
        o.close()
