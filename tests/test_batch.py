"""Batches (many meshes of uneven size in one HIP graph) and launch shapes.

A replayed batch runs the same kernels with the same arguments as the eager
calls, so its results must equal the eager ones bit for bit; each replay is
checked after eager work that overwrote the same frames in between.  Launch
shapes change only the schedule, never a ray's result."""
from __future__ import annotations

import numpy as np
import pytest


def _camera(femrt, hm):
    lo = hm.vertices.reshape(-1, 3).min(0)
    hi = hm.vertices.reshape(-1, 3).max(0)
    return femrt.auto_camera(lo, hi)


def _snapshot(fr, n_iso):
    h = fr.hits()
    out = {k: v.copy() for k, v in h.items()}
    if n_iso:
        for k, v in fr.iso(n_iso).items():
            out["iso_" + k] = v.copy()
    return out


def _same(a, b):
    """Equal hit/root records; t, rst and u are defined only where a hit or
    root exists (elsewhere they keep whatever the frame held before)."""
    assert a.keys() == b.keys()
    np.testing.assert_array_equal(a["elem"], b["elem"], err_msg="elem")
    if "iso_elem" in a:
        np.testing.assert_array_equal(a["iso_elem"], b["iso_elem"], err_msg="iso_elem")
    for k in a:
        if k in ("elem", "iso_elem"):
            continue
        valid = (a["iso_elem"] if k.startswith("iso_") else a["elem"]) >= 0
        x, y = a[k], b[k]
        if x.ndim > valid.ndim:     # rst: (..., 3, n) against (..., n)
            valid = np.broadcast_to(np.expand_dims(valid, -2), x.shape)
        np.testing.assert_array_equal(x[valid], y[valid], err_msg=k)


@pytest.mark.gpu
def test_batch_graph_replays_match_eager(femrt, gpu_ctx):
    ctx = gpu_ctx
    meshes = [femrt.gen_structured_hex(2, 3, 4, pert=1, amp=0.05, seed=1),
              femrt.gen_structured_hex(5, 4, 6, pert=2, amp=0.06, seed=2),
              femrt.gen_structured_hex(3, 3, 3, pert=1, amp=0.04, seed=3)]
    sizes = [(32, 24), (64, 40), (20, 20)]
    guesses = [femrt.GUESS_PLAIN, femrt.GUESS_WARM, femrt.GUESS_LEARNED]
    isos = [[0.0], [-0.1, 0.0, 0.1], [0.05]]
    dm, fr, prm, eager = [], [], [], []
    mlp = femrt.Mlp(ctx, seed=7)   # untrained: the fallbacks still give exact roots
    for i, hm in enumerate(meshes):
        m = femrt.Mesh(ctx, hm)
        m.set_guess_mlp(mlp)
        W, H = sizes[i]
        f = femrt.Frame(ctx, W * H)
        f.camera_rays(_camera(femrt, hm), W, H)
        p = femrt.trace_params(isos[i], guess=guesses[i], flags=1)
        f.trace(m, p)
        eager.append(_snapshot(f, len(isos[i])))
        assert (eager[-1]["elem"] >= 0).sum() > 50
        dm.append(m)
        fr.append(f)
        prm.append(p)
    # a path-traced render of the second mesh in the same graph
    Wp, Hp, spp, depth, seed = 24, 16, 2, 2, 99
    cam12 = femrt.camera_basis((3.0, 2.2, 2.6), (0.5, 0.5, 0.5), (0, 1, 0), 35.0, Wp / Hp)
    mats = femrt.materials(femrt.DEFAULT_MATERIALS)
    fp = femrt.Frame(ctx, Wp * Hp * spp)
    img_eager = fp.render_path(dm[1], cam12, Wp, Hp, spp, depth, mats, seed)

    b = femrt.Batch(ctx)
    for m, f, p in zip(dm, fr, prm):
        b.add_trace(m, f, p)
    b.add_paths(dm[1], fp, cam12, Wp, Hp, spp, depth, mats, seed)
    b.capture()
    other = femrt.trace_params([0.3], guess=femrt.GUESS_PLAIN, flags=1)
    for rep in range(3):
        # eager work between replays overwrites every frame of the batch
        for i, (m, f) in enumerate(zip(dm, fr)):
            f.trace(dm[(i + 1) % len(dm)], other)
        fp.render_path(dm[0], cam12, Wp, Hp, spp, 1, mats, seed + 1 + rep)
        b.launch()
        ctx.sync()
        for i, f in enumerate(fr):
            _same(_snapshot(f, len(isos[i])), eager[i])
        np.testing.assert_array_equal(fp.rgba()[:Wp * Hp], img_eager, err_msg=f"replay {rep}")
    b.close()
    for o in fr + [fp] + dm:
        o.close()
    mlp.close()


@pytest.mark.gpu
def test_batch_jobs_on_one_frame_keep_their_order(femrt, gpu_ctx):
    """The graph runs the jobs of different frames side by side on parallel
    branches; jobs on one frame stay in order.  Six jobs: three frames, two
    traces each, the second trace of a frame overwriting the first."""
    ctx = gpu_ctx
    hms = [femrt.gen_structured_hex(n, 3, 4, pert=1, amp=0.05, seed=10 + n) for n in (2, 3, 4)]
    ms = [femrt.Mesh(ctx, hm) for hm in hms]
    W, H = 40, 30
    frames, eager, jobs = [], [], []
    for i in range(3):
        f = femrt.Frame(ctx, W * H)
        f.camera_rays(_camera(femrt, hms[i]), W, H)
        frames.append(f)
    for i in range(3):
        jobs.append((ms[i], frames[i], femrt.trace_params([0.0], guess=femrt.GUESS_WARM, flags=1)))
    for i in range(3):
        jobs.append((ms[(i + 1) % 3], frames[i], femrt.trace_params([0.1, -0.1], guess=femrt.GUESS_WARM, flags=1)))
    for m, f, p in jobs:          # eager: the last job of each frame wins
        f.trace(m, p)
    eager = [_snapshot(f, 2) for f in frames]
    b = femrt.Batch(ctx)
    for m, f, p in jobs:
        b.add_trace(m, f, p)
    b.capture()
    for rep in range(2):
        for i, f in enumerate(frames):   # overwrite with something else first
            f.trace(ms[i], femrt.trace_params([0.3], guess=femrt.GUESS_PLAIN, flags=1))
        b.launch()
        ctx.sync()
        for i, f in enumerate(frames):
            _same(_snapshot(f, 2), eager[i])
    b.close()
    for o in frames + ms:
        o.close()


@pytest.mark.gpu
def test_batch_rejects_uncapturable_jobs(femrt, gpu_ctx):
    ctx = gpu_ctx
    hm = femrt.gen_structured_hex(2, 3, 3, pert=1, amp=0.04, seed=4)
    m = femrt.Mesh(ctx, hm)
    f = femrt.Frame(ctx, 64)
    b = femrt.Batch(ctx)
    with pytest.raises(femrt.FemrtError):
        b.capture()                                   # no jobs
    with pytest.raises(femrt.FemrtError):
        b.add_trace(m, f, femrt.trace_params([0.0], flags=1 | femrt.TRACE_BINNED_WALK))
    f.counters(True)
    with pytest.raises(femrt.FemrtError):
        b.add_trace(m, f, femrt.trace_params([0.0]))
    f.counters(False)
    b.add_trace(m, f, femrt.trace_params([0.0]))
    b.capture()
    b.launch()
    ctx.sync()
    b.close()
    f.close()
    m.close()


@pytest.mark.gpu
def test_autotune_keeps_results_and_picks_a_candidate(femrt, gpu_ctx):
    ctx = gpu_ctx
    hm = femrt.gen_structured_hex(4, 4, 6, pert=2, amp=0.08, seed=11)
    m = femrt.Mesh(ctx, hm)
    W, H = 64, 48
    f = femrt.Frame(ctx, W * H)
    f.camera_rays(_camera(femrt, hm), W, H)
    p = femrt.trace_params([-0.1, 0.0, 0.1], guess=femrt.GUESS_WARM, flags=1)
    assert m.launch_cfg() == (256, 4, 8, 8, 1)
    f.trace(m, p)
    before = _snapshot(f, 3)
    cfg = m.autotune(f, p, reps=2)
    assert cfg[0] in (64, 128, 256) and cfg[1] in (2, 4, 8, 16) and cfg[2] in (4, 8, 16)
    assert cfg[3] in (4, 8, 16) and cfg[4] in (1, 2, 4)
    assert m.launch_cfg() == cfg
    # every shape, tile and queue split included, gives the same results bit for bit
    for shape in [cfg, (64, 2, 4, 4, 2), (128, 16, 16, 16, 4), (256, 1, 8, 4, 4), (64, 4, 8, 16, 2)]:
        m.launch_cfg(shape)
        f.trace(m, p)
        _same(_snapshot(f, 3), before)
    # the first values alone set only those
    m.launch_cfg((256, 4, 8))
    assert m.launch_cfg() == (256, 4, 8, 16, 2)
    for bad in [(96, 4, 8), (256, 4, 8, 12), (256, 4, 8, 8, 3), (256, 4, 8, 8, 1, 1)]:
        with pytest.raises(femrt.FemrtError):
            m.launch_cfg(bad)
    assert m.launch_cfg() == (256, 4, 8, 16, 2)
    f.close()
    m.close()
