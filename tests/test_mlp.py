"""Learned Newton guess: the FP64 network oracle pinned by finite differences
and exact bfloat16 rounding (CPU), and the GPU's BF16-MFMA forward, backward,
training and sample generation against it."""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle.mlp_oracle as mo


def _data(seed, n, scale=1.0):
    rng = np.random.default_rng(seed)
    feats = rng.uniform(-1.2, 1.2, (n, mo.N_IN)) * scale
    tgt = rng.uniform(-1, 1, (n, 3))
    return feats, tgt


# ------------------------------------------------------------------ oracle

def test_bf16_rounding_is_nearest_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-3, 500), rng.normal(0, 1e3, 500)])
    x = x.astype(np.float32).astype(np.float64)
    got = mo.bf16(x)
    for v, g in zip(x, got):
        if v == 0:
            assert g == 0
            continue
        e = np.floor(np.log2(abs(v)))
        ulp = 2.0 ** (e - 7)                # 8 significant bits
        lo = np.floor(v / ulp) * ulp
        cands = [lo, lo + ulp]
        best = min(cands, key=lambda c: (abs(c - v), (round(c / ulp) % 2)))
        assert g == best, (v, g, best)
    # ties go to the even significand
    assert mo.bf16(1.0 + 2.0 ** -8) == 1.0
    assert mo.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6


def test_oracle_gradient_matches_finite_differences():
    p = mo.glorot_params(1)
    feats, tgt = _data(2, 48)
    loss, g = mo.loss_grad(p, feats, tgt, emulate=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(np.arange(mo.OFF_W1, mo.OFF_W2), 12, replace=False),
                          rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 12, replace=False),
                          np.arange(mo.OFF_B2, mo.OFF_B2 + 4),
                          rng.choice(np.arange(mo.OFF_W3, mo.OFF_B3), 8, replace=False),
                          np.arange(mo.OFF_B3, mo.N_PARAM),
                          np.arange(27 * 64, 27 * 64 + 3)])     # first-layer bias row
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_grad(pp, feats, tgt, False)[0] - mo.loss_grad(pm, feats, tgt, False)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_oracle_emulation_is_close_to_exact():
    p = mo.glorot_params(4)
    feats, tgt = _data(5, 256)
    y_e, _ = mo.forward(p, feats, True)
    y_x, _ = mo.forward(p, feats, False)
    assert np.max(np.abs(y_e - y_x)) < 0.05
    l_e, g_e = mo.loss_grad(p, feats, tgt, True)
    l_x, g_x = mo.loss_grad(p, feats, tgt, False)
    assert l_e == pytest.approx(l_x, rel=0.05)
    assert np.dot(g_e, g_x) / (np.linalg.norm(g_e) * np.linalg.norm(g_x)) > 0.999


def test_bracket_oracle_gradient_matches_finite_differences():
    """The bracket network (explicit first-layer bias, two outputs, pinball
    loss at quantiles 0.05 and 0.95): FD of the exact FP64 loss, away from
    the pinball kinks (continuous random targets put no residual at 0)."""
    spec = mo.BRACKET
    p = mo.glorot_params(11, spec=spec)
    rng = np.random.default_rng(12)
    feats = rng.uniform(-1.2, 1.2, (48, spec.n_in))
    t = rng.uniform(-0.5, 1.5, 48)
    tgt = np.stack([t, t], axis=1)
    loss, g = mo.loss_grad(p, feats, tgt, emulate=False, spec=spec)
    o = spec.offsets
    idx = np.concatenate([rng.choice(np.arange(o["W1"], o["b1"]), 10, replace=False),
                          np.arange(o["b1"], o["b1"] + 6),
                          rng.choice(np.arange(o["W2"], o["b2"]), 10, replace=False),
                          np.arange(o["b2"], o["b2"] + 4),
                          rng.choice(np.arange(o["W3"], o["b3"]), 8, replace=False),
                          np.arange(o["b3"], o["end"])])
    h = 1e-7
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_grad(pp, feats, tgt, False, spec)[0] - mo.loss_grad(pm, feats, tgt, False, spec)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-4, abs=1e-9), k


def test_pinball_loss_closed_form():
    """Pinball loss of a constant prediction: minimised at the quantile of
    the targets (brute force over candidate constants)."""
    spec = mo.BRACKET
    t = np.random.default_rng(3).normal(0, 1, 4001)
    cands = np.linspace(-3, 3, 1201)
    for j, tau in enumerate(spec.quantiles):
        losses = []
        for c in cands:
            y = np.zeros((t.size, 2))
            y[:, j] = c
            losses.append(mo.loss_and_output_grad(y, np.stack([t, t], 1), t.size, spec)[0])
        best = cands[int(np.argmin(losses))]
        assert abs(best - np.quantile(t, tau)) < 0.02


# -------------------------------------------------------------------- GPU

def _blocks(g):
    return dict(W1=g[mo.OFF_W1:mo.OFF_W2], W2=g[mo.OFF_W2:mo.OFF_B2], b2=g[mo.OFF_B2:mo.OFF_W3],
                W3=g[mo.OFF_W3:mo.OFF_B3], b3=g[mo.OFF_B3:mo.N_PARAM])


@pytest.mark.gpu
def test_mlp_forward_matches_oracle(femrt, gpu_ctx):
    mlp = femrt.Mlp(gpu_ctx, seed=1)
    p = mo.glorot_params(6).astype(np.float32)
    mlp.set_params(p)
    np.testing.assert_array_equal(mlp.params(), p)
    for n in (1, 63, 64, 65, 1000):   # ragged sizes: partial waves and blocks
        feats, _ = _data(7 + n, n)
        feats = feats.astype(np.float32)
        y = mlp.forward(feats)
        y_e, _ = mo.forward(p, feats, True)
        y_x, _ = mo.forward(p, feats, False)
        # same rounding points: differences come from FP32 accumulation order
        # and the tanh evaluation flipping an occasional bf16 rounding
        assert np.max(np.abs(y - y_e)) < 3e-3, n
        assert np.mean(np.abs(y - y_e)) < 2e-4, n
        assert np.max(np.abs(y - y_x)) < 0.05, n
    mlp.close()


@pytest.mark.gpu
def test_mlp_gradient_matches_oracle(femrt, gpu_ctx):
    mlp = femrt.Mlp(gpu_ctx, seed=2)
    p = mo.glorot_params(8).astype(np.float32)
    mlp.set_params(p)
    for n in (64, 200, 4096):
        feats, tgt = _data(9 + n, n)
        feats, tgt = feats.astype(np.float32), tgt.astype(np.float32)
        loss, g = mlp.loss_grad(feats, tgt)
        l_e, g_e = mo.loss_grad(p, feats, tgt, True)
        l_x, g_x = mo.loss_grad(p, feats, tgt, False)
        assert loss == pytest.approx(l_e, rel=1e-4), n
        assert loss == pytest.approx(l_x, rel=0.05), n
        gb, eb, xb = _blocks(g), _blocks(g_e), _blocks(g_x)
        for k in gb:
            err = np.linalg.norm(gb[k] - eb[k]) / np.linalg.norm(eb[k])
            assert err < 1e-2, (n, k, err)
            cos = np.dot(gb[k], xb[k]) / (np.linalg.norm(gb[k]) * np.linalg.norm(xb[k]))
            assert cos > 0.995, (n, k, cos)
    mlp.close()


def _small_mesh(femrt, gpu_ctx, n=4, seed=31):
    hm = femrt.gen_structured_hex(n, 4, 3, pert=2, amp=0.08, seed=seed)
    return hm, femrt.Mesh(gpu_ctx, hm)


@pytest.mark.gpu
def test_training_samples_are_newton_preimages(femrt, oracle, gpu_ctx):
    hm, mesh = _small_mesh(femrt, gpu_ctx, n=3)
    feats, tgt = femrt.mlp_samples(mesh, 300, seed=5)
    boxes = mesh.boxes()[:, 0, :]
    lo, hi = boxes[:, :3].astype(np.float64), boxes[:, 3:].astype(np.float64)
    norm_corners = 2 * (hm.vertices.astype(np.float64) - lo[:, None, :]) / (hi - lo)[:, None, :] - 1
    orc = oracle.Oracle.from_mesh(hm)
    assert np.all(np.abs(tgt) <= 1.001)
    for i in range(feats.shape[0]):
        c = feats[i, 3:].reshape(8, 3)
        e = int(np.argmin(np.abs(norm_corners - c[None]).reshape(hm.n_elem, -1).max(1)))
        assert np.max(np.abs(norm_corners[e] - c)) < 1e-5
        x = lo[e] + (feats[i, :3].astype(np.float64) + 1) * 0.5 * (hi[e] - lo[e])
        xm, _ = orc.map(e, tgt[i].astype(np.float64))
        np.testing.assert_allclose(xm, x, atol=2e-6)
    mesh.close()


@pytest.mark.gpu
def test_training_reduces_loss_and_guess_error(femrt, gpu_ctx):
    hm, mesh = _small_mesh(femrt, gpu_ctx)
    mlp = femrt.Mlp(gpu_ctx, seed=3)
    feats, tgt = femrt.mlp_samples(mesh, 4096, seed=999)  # held out
    before = np.abs(mlp.forward(feats) - tgt)
    losses = mlp.train(mesh, n_samples=1 << 17, batch=2048, epochs=6, lr=3e-3, lr_final=1e-4,
                       seed=7)
    after = np.abs(mlp.forward(feats) - tgt)
    print("epoch losses", losses, "median err before/after", np.median(before), np.median(after))
    assert losses[-1] < 0.05 * losses[0]
    assert np.median(after) < 0.03
    assert np.percentile(after.max(1), 99) < 0.2
    assert np.median(after) < 0.1 * np.median(before)
    mlp.close()
    mesh.close()


@pytest.mark.gpu
def test_training_is_reproducible(femrt, gpu_ctx):
    """Each training block writes its partial gradient and the Adam step sums
    them in block order (no floating-point atomics), so the same seeds give
    the same network bit for bit, and the same gradient."""
    hm, mesh = _small_mesh(femrt, gpu_ctx)
    out = []
    for _ in range(2):
        mlp = femrt.Mlp(gpu_ctx, seed=11)
        losses = mlp.train(mesh, n_samples=1 << 15, batch=2048, epochs=2, lr=3e-3, lr_final=1e-4, seed=5)
        feats, tgt = femrt.mlp_samples(mesh, 1000, seed=17)
        out.append((mlp.params(), np.asarray(losses), mlp.loss_grad(feats, tgt)))
        mlp.close()
    (p0, l0, (lg0, g0)), (p1, l1, (lg1, g1)) = out
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(l0, l1)
    assert lg0 == lg1
    np.testing.assert_array_equal(g0, g1)
    mesh.close()


@pytest.mark.gpu
def test_learned_guess_keeps_parity_and_what_it_gains(femrt, oracle, gpu_ctx):
    """Every walk form with network guesses gives the oracle's results.  What
    the trained network gains:

    * over the plain guess rst = 0 (the network with all-zero parameters):
      fewer point-inversion iterations when it guesses every sample, and when
      it guesses each segment's first sample of the sequential scan;
    * over the engineered warm starts of the same walk: nothing measurable.
      In the sequential scan the network's first-sample guess (median error
      about 0.05 in reference units) is no better than the warm start from
      the entry point, and the cooperative scan (warm guesses only) starts
      every sample from the cubic Hermite interpolant of rst(t) (error
      O(dt^4)).  The test asserts parity and that the network costs no
      iterations over warm starts in the sequential scan, and prints the
      ratios.
    These comparisons switch off the default Hermite start of each
    segment's first sample (FEMRT_TRACE_NO_HERMITE_START), which takes the
    network's place wherever the predicted exit left the tangents; the
    default walk (form "hermite") keeps parity and needs no more iterations
    than the same walk without it.
    Trained on one seeded mesh, evaluated on another."""
    import oracle.parity as parity
    train_hm = femrt.gen_structured_hex(6, 4, 6, pert=2, amp=0.08, seed=1)
    train_mesh = femrt.Mesh(gpu_ctx, train_hm)
    mlp = femrt.Mlp(gpu_ctx, seed=4)
    losses = mlp.train(train_mesh, n_samples=1 << 18, batch=4096, epochs=6, lr=3e-3,
                       lr_final=1e-4, seed=11)
    print("losses", losses)
    zero = femrt.Mlp(gpu_ctx, seed=5)
    zero.set_params(np.zeros(zero.n_param, np.float32))

    hm = femrt.gen_structured_hex(4, 4, 6, pert=2, amp=0.08, seed=99)
    mesh = femrt.Mesh(gpu_ctx, hm)
    W, H = 64, 48
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 1])
    fr = femrt.Frame(gpu_ctx, W * H)
    fr.camera_rays(cam, W, H)
    isos = [-0.1, 0.0, 0.1]
    og, oiso = oracle.Oracle.from_mesh(hm).trace(fr.get_rays(), isos)
    nh = femrt.TRACE_NO_HERMITE_START
    forms = {"every": 1 | femrt.TRACE_LEARNED_EVERY_SAMPLE | nh, "seq": 1 | nh, "hermite": 1,
             "default": 1 | femrt.TRACE_COOP_SCAN}
    work = {}
    for form, flags in forms.items():
        # the cooperative scan (samples spread over the lanes, Hermite starts)
        # takes warm guesses only
        runs = [("warm", femrt.GUESS_WARM, mlp)]
        if form in ("every", "seq"):
            runs += [("trained", femrt.GUESS_LEARNED, mlp), ("zero", femrt.GUESS_LEARNED, zero)]
        elif form == "hermite":
            runs += [("trained", femrt.GUESS_LEARNED, mlp)]
        for name, guess, net in runs:
            mesh.set_guess_mlp(net)
            fr.counters(True)
            fr.trace(mesh, femrt.trace_params(isos, guess=guess, flags=flags))
            rep = parity.compare(fr.hits(), fr.iso(len(isos)), og, oiso, isos)
            work[name, form] = fr.counters()
            print(form, name, json.dumps(rep), {k: work[name, form][k] for k in ("invert_it", "mlp_infer")})
            assert rep["ok"], (form, name, rep)
    it = {k: v["invert_it"] for k, v in work.items()}
    print(json.dumps({"learned_over_warm_iterations": {"seq": it["trained", "seq"] / it["warm", "seq"],
                                                       "seq_vs_default_walk": it["trained", "seq"] / it["warm", "default"]},
                      "learned_over_zero_net_iterations": {f: it["trained", f] / it["zero", f] for f in ("every", "seq")}}))
    assert work["trained", "every"]["mlp_infer"] > 0 and work["trained", "seq"]["mlp_infer"] > 0
    assert it["trained", "every"] < 0.8 * it["zero", "every"]
    assert it["trained", "seq"] < it["zero", "seq"]
    assert work["trained", "seq"]["mlp_infer"] < 0.25 * work["trained", "every"]["mlp_infer"]
    assert it["trained", "seq"] <= 1.05 * it["warm", "seq"]
    print(json.dumps({"hermite_over_entry_start_iterations": it["warm", "hermite"] / it["warm", "seq"]}))
    assert it["warm", "hermite"] <= it["warm", "seq"]
    assert it["trained", "hermite"] <= it["trained", "seq"]
    with pytest.raises(femrt.FemrtError):
        mesh.set_guess_mlp(None)
        fr.trace(mesh, femrt.trace_params(isos, guess=femrt.GUESS_LEARNED))
    for o in (fr, mesh, train_mesh, mlp, zero):
        o.close()
