"""Parity at full size on the GPU for the three single-frame workloads:

* 32^3 curved hex, Q=4, amplitude 0.08 h, P=6, 1920x1080, isovalues
  {-0.1, 0, 0.1} (the benchmark frame);
* 64^3 curved hex, Q=4, face bubbles of peak 0.15 h, P=6, 3840x2160
  (learned guesses on a larger, more curved mesh);
* 64^3 curved hex, Q=4, face bubbles of peak 0.15 h, P=8 multi-frequency field with
  four isovalues, 1920x1080 at 16 jittered samples per pixel.

Each frame is traced whole on the GPU.  A seeded random sample of hit and
miss rays is checked against the FP64 oracle at the acceptance bar (the
oracle would take hours over the whole frame), and the whole frame is
checked by a size-independent property: the learned, warm-start and plain
Newton variants start their FP32 iterations from different points but must
converge to the same FP64 results on every ray, ties aside; so must the
scan that evaluates every sample (no certified skipping).  Every ray where
two variants disagree is then traced by the oracle too, and both variants
must meet the bar there (ties aside); so must every ray whose walk stalled,
and the 64^3 P=8 frame's rays whose walks once ended stalled.  (On that
frame the FP64 oracle needs seconds per ray, so its disagreements are
checked on an even subsample of 240.)
"""
from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_SAMPLE_HIT, N_SAMPLE_MISS = 240, 30
# GPU-vs-GPU agreement over the whole frame: the same FP64 root to 1e-9
# relative; a ray may differ only where an FP32 sign or id decision is a tie
# (a sample of u - iso within about 1e-5 of zero, or two faces within 1e-6 t)
SAME_T_REL = 1e-9
MAX_GEOM_DIFF_FRAC = 1e-4
MAX_ISO_DIFF_FRAC = 2e-3


def _sample_check(oracle, hm, fr, isos, seed, extra=()):
    import oracle.parity as parity
    gh, giso, rays = fr.hits(), fr.iso(len(isos)), fr.get_rays()
    rng = np.random.default_rng(seed)
    hit = np.flatnonzero(gh["elem"] >= 0)
    miss = np.flatnonzero(gh["elem"] < 0)
    sel = np.concatenate([rng.choice(hit, min(N_SAMPLE_HIT, hit.size), replace=False),
                          rng.choice(miss, min(N_SAMPLE_MISS, miss.size), replace=False)])
    sel = np.union1d(sel, np.asarray(extra, np.int64))

    def sub(d):
        return {k: (v[..., sel] if isinstance(v, np.ndarray) else v) for k, v in d.items()}
    og, oiso = oracle.Oracle.from_mesh(hm).trace(np.ascontiguousarray(rays[:, sel]), isos)
    rep = parity.compare(sub(gh), sub(giso), og, oiso, isos)
    return rep, gh, giso


MAX_ORACLE_DISAGREE = 2000


def _disagreement_vs_oracle(oracle, hm, rays, isos, traces, ids, cap=MAX_ORACLE_DISAGREE):
    """Every ray where two GPU traces disagree (at most `cap` of them, an even
    subsample beyond that) against the oracle: each trace must meet the bar
    there, ties aside (a tie is where the FP32 decision is not determined:
    two faces within 1e-6 t, a face coordinate at the acceptance bound, or a
    scan sample of u - iso within the oracle's tie margin of zero)."""
    import oracle.parity as parity
    ids = np.asarray(ids, np.int64)
    total = int(ids.size)
    if ids.size > cap:
        ids = ids[np.linspace(0, ids.size - 1, cap).astype(np.int64)]
    if ids.size == 0:
        return {"disagreeing_rays": 0}
    orc = oracle.Oracle.from_mesh(hm)
    parts = []
    for c0 in range(0, ids.size, 200):   # in chunks, reporting progress
        parts.append(orc.trace(np.ascontiguousarray(rays[:, ids[c0:c0 + 200]]), isos))
        print(json.dumps({"oracle_rays_done": min(c0 + 200, int(ids.size)), "of": int(ids.size)}), flush=True)
    og = {k: np.concatenate([p[0][k] for p in parts], axis=-1) for k in parts[0][0]}
    oi = {k: np.concatenate([p[1][k] for p in parts], axis=-1) for k in parts[0][1]}
    out = {"disagreeing_rays": total, "checked": int(ids.size)}
    for name, (gh, giso) in traces.items():
        sub = lambda d: {k: (v[..., ids] if isinstance(v, np.ndarray) else v) for k, v in d.items()}  # noqa: E731
        rep = parity.compare(sub(gh), sub(giso), og, oi, isos)
        out[name] = {"ok": rep["ok"], "ties": rep["ties"], "hit_miss_mismatch": rep["hit_miss_mismatch"],
                     "id_mismatch": rep["id_mismatch"],
                     "iso": [(r["found_mismatch"], r["elem_mismatch"], r["t_rel_bad_nontie"], r["ties"])
                             for r in rep["iso"]]}
    return out


def _disagreeing_rays(a_hits, a_iso, b_hits, b_iso):
    ha, hb = a_hits["elem"] >= 0, b_hits["elem"] >= 0
    both = ha & hb
    trel = np.abs(a_hits["t"] - b_hits["t"]) / np.maximum(np.abs(b_hits["t"]), 1e-300)
    bad = (ha != hb) | (both & ((a_hits["elem"] != b_hits["elem"]) | (a_hits["face"] != b_hits["face"]) |
                                (trel > SAME_T_REL)))
    for k in range(a_iso["elem"].shape[0]):
        fa, fb = a_iso["elem"][k] >= 0, b_iso["elem"][k] >= 0
        tr = np.abs(a_iso["t"][k] - b_iso["t"][k]) / np.maximum(np.abs(b_iso["t"][k]), 1e-300)
        bad |= (fa != fb) | (fa & fb & ((a_iso["elem"][k] != b_iso["elem"][k]) | (tr > SAME_T_REL)))
    return np.flatnonzero(bad)


def _frame_agreement(a_hits, a_iso, b_hits, b_iso):
    """Whole-frame comparison of two GPU traces of the same rays."""
    ha, hb = a_hits["elem"] >= 0, b_hits["elem"] >= 0
    n_hit = max(1, int(ha.sum()))
    both = ha & hb
    same = both & (a_hits["elem"] == b_hits["elem"]) & (a_hits["face"] == b_hits["face"])
    trel = np.abs(a_hits["t"] - b_hits["t"]) / np.maximum(np.abs(b_hits["t"]), 1e-300)
    geom_diff = int((ha != hb).sum() + (both & ~same).sum() + (same & (trel > SAME_T_REL)).sum())
    out = {"hits": int(ha.sum()), "geom_diff": geom_diff, "geom_diff_frac": geom_diff / n_hit,
           "t_rel_max_same": float(trel[same].max()) if same.any() else 0.0, "iso": []}
    for k in range(a_iso["elem"].shape[0]):
        fa, fb = a_iso["elem"][k] >= 0, b_iso["elem"][k] >= 0
        b = fa & fb
        tr = np.abs(a_iso["t"][k] - b_iso["t"][k]) / np.maximum(np.abs(b_iso["t"][k]), 1e-300)
        se = b & (a_iso["elem"][k] == b_iso["elem"][k]) & (tr <= SAME_T_REL)
        diff = int((fa != fb).sum() + (b & ~se).sum())
        out["iso"].append({"found": int(fa.sum()), "diff": diff,
                           "diff_frac": diff / max(1, int(fa.sum()))})
    return out


def _trained_mlp(femrt, ctx, pert, amp, seed):
    """Guess network trained by the library's kernels on a 16^3 mesh of the
    same family (another seed)."""
    thm = femrt.gen_structured_hex(16, 4, 2, pert=pert, amp=amp, seed=seed)
    tmesh = femrt.Mesh(ctx, thm, build_bvh=False)
    mlp = femrt.Mlp(ctx, seed=seed)
    losses = mlp.train(tmesh, n_samples=1 << 20, batch=4096, epochs=3, lr=3e-3, lr_final=1e-4,
                       seed=seed + 1)
    tmesh.close()
    assert losses[-1] < 0.5 * losses[0]
    return mlp


def _run(femrt, oracle, ctx, hm, rays_fn, n_rays, isos, pert, amp, seed, variants, extra=(),
         oracle_cap=MAX_ORACLE_DISAGREE):
    mesh = femrt.Mesh(ctx, hm)
    mlp = _trained_mlp(femrt, ctx, pert, amp, seed)
    mesh.set_guess_mlp(mlp)
    fr = femrt.Frame(ctx, n_rays)
    rays_fn(fr)
    fr.counters(True)
    fr.trace(mesh, femrt.trace_params(isos, guess=femrt.GUESS_LEARNED))
    cnt = fr.counters(False)
    n_st, st_ids = fr.stalled_rays()
    rep, gh, giso = _sample_check(oracle, hm, fr, isos, seed, np.union1d(np.asarray(extra, np.int64), st_ids))
    print(json.dumps({"learned_vs_oracle": rep, "counters": {k: cnt[k] for k in ("hit_rays", "mlp_infer",
                                                                                     "brackets", "walks_stalled",
                                                                                     "walk_skips", "walk_jumps")}}),
          flush=True)
    assert rep["ok"], rep
    assert rep["oracle_hits"] >= N_SAMPLE_HIT - 5
    assert cnt["walks_stalled"] == 0
    rays = fr.get_rays()
    for name in variants:
        guess, flags = name if isinstance(name, tuple) else (name, 1)
        fr.trace(mesh, femrt.trace_params(isos, guess=guess, flags=flags))
        vh, vi = fr.hits(), fr.iso(len(isos))
        agree = _frame_agreement(gh, giso, vh, vi)
        dis = _disagreement_vs_oracle(oracle, hm, rays, isos, {"learned": (gh, giso), "variant": (vh, vi)},
                                      _disagreeing_rays(gh, giso, vh, vi), oracle_cap)
        print(json.dumps({"learned_vs_variant": [guess, flags], **agree, "disagreements_vs_oracle": dis}),
              flush=True)
        assert agree["geom_diff_frac"] <= MAX_GEOM_DIFF_FRAC, agree
        assert all(r["diff_frac"] <= MAX_ISO_DIFF_FRAC for r in agree["iso"]), agree
        # the plain variant's face roots are the reference's nine seeds, which
        # miss the nearer crossing of a face crossed twice: reported, not gated
        gated = ("learned",) if guess == femrt.GUESS_PLAIN else ("learned", "variant")
        assert all(dis[k]["ok"] for k in gated if k in dis), dis
    for o in (fr, mesh, mlp):
        o.close()
    return rep, cnt


def test_bench_frame_32cube_p6_1080p(femrt, oracle, gpu_ctx):
    hm = femrt.gen_structured_hex(32, 4, 6, pert=2, amp=0.08, seed=20240611)
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 1])
    W, H = 1920, 1080
    isos = [-0.1, 0.0, 0.1]
    rep, cnt = _run(femrt, oracle, gpu_ctx, hm, lambda fr: fr.camera_rays(cam, W, H), W * H, isos,
                    2, 0.08, 101, [femrt.GUESS_WARM, (femrt.GUESS_WARM, 1 | femrt.TRACE_COOP_SCAN),
                                   femrt.GUESS_PLAIN, (femrt.GUESS_LEARNED, 1 | femrt.TRACE_NO_SAMPLE_SKIP)])
    assert all(r["oracle_found"] > 100 for r in rep["iso"])


def test_learned_64cube_amp015_4k(femrt, oracle, gpu_ctx):
    hm = femrt.gen_structured_hex(64, 4, 6, pert=3, amp=0.15, seed=404)
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 1])
    W, H = 3840, 2160
    isos = [0.0]
    rep, cnt = _run(femrt, oracle, gpu_ctx, hm, lambda fr: fr.camera_rays(cam, W, H), W * H, isos,
                    3, 0.15, 202, [femrt.GUESS_WARM])
    assert cnt["hit_rays"] > 0.3 * W * H
    assert rep["iso"][0]["oracle_found"] > 150


def test_multifrequency_64cube_p8_four_isos_16spp(femrt, oracle, gpu_ctx):
    base = femrt.gen_structured_hex(64, 4, 2, pert=3, amp=0.15, seed=505)
    hm = femrt.gen_field_waves(base, 8, n_waves=8, seed=506)
    assert hm.nf == 9
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 1])
    W, H, spp = 1920, 1080, 16
    cam12 = femrt.camera_basis(cam[:3], cam[3:6], cam[6:9], float(cam[9]), W / H)
    isos = [-0.6, -0.2, 0.2, 0.6]
    rep, cnt = _run(femrt, oracle, gpu_ctx, hm, lambda fr: fr.camera_rays_jitter(cam12, W, H, spp, 507),
                    W * H * spp, isos, 3, 0.15, 303, [femrt.GUESS_WARM],
                    extra=_stalled_fixture("p8_1920x1080x16"), oracle_cap=240)
    # most segments' field bounds exclude every open isovalue: runs of them
    # are passed by BVH queries
    assert cnt["walk_jumps"] > 0
    assert all(r["oracle_found"] > 60 for r in rep["iso"])


def _stalled_fixture(name):
    import pathlib
    d = json.loads((pathlib.Path(__file__).parent / "golden" / "stalled_walk_rays.json").read_text())
    return np.array(d[name]["ids"], np.int64)


def test_formerly_stalled_walks_finish_and_match_oracle(femrt, oracle, gpu_ctx):
    """Regression: on this 64^3 P=8 frame eight walks used to end as
    stalled: the rays run through element edges or inside faces two
    elements share, where FP32 face roots cannot resolve the exit and the
    marching fallback finds it (a walk that ran into 16 marched exits in
    total ended there; one ray, inside a shared face, was handed back and
    forth between two elements).  Now only an unbroken run of 16 marched
    exits counts, a stall is skipped past in the element, and no walk ends
    stalled; the rays' crossings match the oracle (whose segments between
    face roots are classified by their midpoints, see test_oracle_brute)."""
    import time
    import oracle.parity as parity
    base = femrt.gen_structured_hex(64, 4, 2, pert=3, amp=0.15, seed=20240611)
    hm = femrt.gen_field_waves(base, 8, n_waves=8, seed=20240611 + 4)
    mesh = femrt.Mesh(gpu_ctx, hm)
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 1])
    W, H, spp = 960, 540, 16
    cam12 = femrt.camera_basis(cam[:3], cam[3:6], cam[6:9], float(cam[9]), W / H)
    fr = femrt.Frame(gpu_ctx, W * H * spp)
    isos = [-0.6, -0.2, 0.2, 0.6]
    fr.camera_rays_jitter(cam12, W, H, spp, 7)
    fr.counters(True)
    gpu_ctx.sync()
    t0 = time.perf_counter()
    fr.trace(mesh, femrt.trace_params(isos, guess=femrt.GUESS_WARM))
    gpu_ctx.sync()
    dt = time.perf_counter() - t0
    cnt = fr.counters(False)
    n_st, ids = fr.stalled_rays()
    print(json.dumps({"s": dt, "walks_stalled": cnt["walks_stalled"], "walk_skips": cnt["walk_skips"],
                      "skipping_rays": int(n_st), "exit_march": cnt["exit_march"]}), flush=True)
    assert cnt["walks_stalled"] == 0
    assert dt < 15.0
    sel = np.union1d(_stalled_fixture("regression_960x540x16"), ids)
    gh, giso, rays = fr.hits(), fr.iso(len(isos)), fr.get_rays()
    og, oi = oracle.Oracle.from_mesh(hm).trace(np.ascontiguousarray(rays[:, sel]), isos)
    sub = lambda d: {k: (v[..., sel] if isinstance(v, np.ndarray) else v) for k, v in d.items()}  # noqa: E731
    rep = parity.compare(sub(gh), sub(giso), og, oi, isos)
    print(json.dumps({"formerly_stalled_vs_oracle": rep}), flush=True)
    assert rep["ok"], rep
    fr.close()
    mesh.close()


@pytest.mark.parametrize("n,amp,seed,W,H", [(32, 0.08, 20240611, 1920, 1080), (64, 0.15, 404, 1280, 720)])
def test_flat_first_hit_equals_nested(femrt, gpu_ctx, n, amp, seed, W, H):
    """The first hit with its face-Newton work flattened across the wave
    (the default for warm and learned guesses) gives every ray the same hit,
    bit for bit, as each lane's search run to completion
    (FEMRT_TRACE_NESTED_FIRST_HIT): the same jobs in the same order per lane,
    culled by the same nearest root."""
    hm = femrt.gen_structured_hex(n, 4, 6, pert=2 if n == 32 else 3, amp=amp, seed=seed)
    mesh = femrt.Mesh(gpu_ctx, hm)
    fr = femrt.Frame(gpu_ctx, W * H)
    fr.camera_rays(femrt.auto_camera([0, 0, 0], [1, 1, 1]), W, H)
    out = []
    for flags in (1, 1 | femrt.TRACE_NESTED_FIRST_HIT):
        fr.counters(True)
        fr.trace(mesh, femrt.trace_params([], guess=femrt.GUESS_WARM, iso_enable=False, flags=flags))
        cnt = fr.counters(False)
        out.append((fr.hits(), cnt))
    (a, ca), (b, cb) = out
    print(json.dumps({"hits": int((a["elem"] >= 0).sum()), "face_newton_it": [ca["first_hit_face_it"],
                                                                             cb["first_hit_face_it"]]}), flush=True)
    assert int((a["elem"] >= 0).sum()) > 0.2 * W * H
    np.testing.assert_array_equal(a["elem"], b["elem"])
    np.testing.assert_array_equal(a["face"], b["face"])
    h = a["elem"] >= 0
    np.testing.assert_array_equal(a["t"][h], b["t"][h])
    np.testing.assert_array_equal(a["rst"][:, h], b["rst"][:, h])
    np.testing.assert_array_equal(a["u"][h], b["u"][h])
    # the same work: face-Newton iterations and slab tests per lane
    assert ca["first_hit_face_it"] == cb["first_hit_face_it"]
    assert ca["first_hit_slab_tests"] == cb["first_hit_slab_tests"]
    fr.close()
    mesh.close()


def test_walk_starts_at_the_polished_hit(femrt, oracle, gpu_ctx):
    """Regression: on the 64^3 P=8 frame these two rays' FP32 first hit is a
    near miss just outside face 0 of the hit element (t 3.99988 and 4.05179);
    the FP64 polish moves on to the face's real crossing (t 4.000034 and
    4.052010).  The walk used to start at the FP32 point, outside the
    element, and missed a crossing 2.5e-3 and 3.3e-4 past the entry; it now
    starts at the polished root."""
    import oracle.parity as parity
    base = femrt.gen_structured_hex(64, 4, 2, pert=3, amp=0.15, seed=505)
    hm = femrt.gen_field_waves(base, 8, n_waves=8, seed=506)
    isos = [-0.6, -0.2, 0.2, 0.6]
    o = [3.0, -2.0, -1.5]
    dirs = [[-0.6632617115974426, 0.6477730870246887, 0.37479862570762634],
            [-0.5812870860099792, 0.7246705293655396, 0.37007826566696167]]
    rays = np.ascontiguousarray(np.array([[o[0]] * 2, [o[1]] * 2, [o[2]] * 2,
                                          [dd[0] for dd in dirs], [dd[1] for dd in dirs], [dd[2] for dd in dirs]]),
                                np.float32)
    mesh = femrt.Mesh(gpu_ctx, hm)
    fr = femrt.Frame(gpu_ctx, 2)
    fr.set_rays(rays)
    og, oi = oracle.Oracle.from_mesh(hm).trace(rays, isos)
    assert oi["elem"][1][0] == 2390 and oi["elem"][2][1] == 3817
    for guess, flags in ((femrt.GUESS_WARM, 1), (femrt.GUESS_WARM, 1 | femrt.TRACE_NESTED_FIRST_HIT),
                         (femrt.GUESS_WARM, 1 | femrt.TRACE_NO_SAMPLE_SKIP),
                         (femrt.GUESS_WARM, 1 | femrt.TRACE_COOP_SCAN)):
        fr.trace(mesh, femrt.trace_params(isos, guess=guess, flags=flags))
        rep = parity.compare(fr.hits(), fr.iso(len(isos)), og, oi, isos)
        print(json.dumps({"guess": guess, "flags": flags, "ok": rep["ok"],
                          "iso_found": [r["gpu_found"] for r in rep["iso"]]}), flush=True)
        assert rep["ok"], rep
    fr.close()
    mesh.close()
