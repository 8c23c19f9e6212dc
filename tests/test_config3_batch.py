"""The batched workload at full size: 64 seeded curved-hex meshes of 8^3 to
32^3 hexes (512 to 32768), Q=3, field orders P in 2..8, each rendered
512x512 at 8 jittered samples per pixel with diffuse bounces to depth 3,
all recorded into one HIP graph (the workload bench.py --workload batch
times, with the same seeds).

Per mesh the eager render runs depth by depth (camera rays, geometry trace,
scatter, ..., resolve).  On a seeded sample of pixels (all 8 of their
paths) every depth is checked against the FP64 oracle: the hit ids and
distance of the traced rays (ties aside), the scattered rays, and in the
end the resolved pixel colour.  Then the graph of all 64 renders is
replayed twice, each time after eager work overwrote every frame, and every
mesh's replayed image must equal its eager image bit for bit.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle.path_oracle as po

pytestmark = pytest.mark.gpu

SEED = 20240611          # the benchmark's seed family
N_MESH, W, H, SPP, DEPTH = 64, 512, 512, 8, 3
Q, AMP, PERT = 3, 0.08, 2
PIXELS_PER_MESH = 2


def _batch_meshes(femrt):
    rng = np.random.default_rng(SEED + 7)
    for i in range(N_MESH):
        n = int(rng.integers(8, 33))
        p = int(rng.integers(2, 9))
        yield i, n, p, femrt.gen_structured_hex(n, Q, p, pert=PERT, amp=AMP, seed=SEED + 100 + i)


def _eager_with_oracle(femrt, oracle, parity, mesh, hm, fr, cam12, mats, specs, seed, rng):
    """Depth-by-depth eager render; returns the image and a report of the
    oracle checks on the sampled pixels' paths."""
    fr.camera_rays_jitter(cam12, W, H, SPP, seed)
    fr.path_begin()
    prm = femrt.trace_params([], guess=femrt.GUESS_WARM, iso_enable=False, flags=0)
    orc = oracle.Oracle.from_mesh(hm)
    rep = {"traced": 0, "hits": 0, "ties": 0, "t_rel_max": 0.0, "dir_err_max": 0.0}
    ids = None
    for depth in range(DEPTH):
        rays_all = fr.get_rays()
        fr.trace(mesh, prm)
        gh_all = fr.hits()
        if ids is None:
            # pixels whose first path hits the mesh, plus one random pixel
            first = gh_all["elem"][::SPP] >= 0
            cand = np.flatnonzero(first)
            pix = rng.choice(cand, min(PIXELS_PER_MESH - 1, cand.size), replace=False)
            pix = np.union1d(pix, rng.integers(0, W * H, 1))
            ids = (pix[:, None] * SPP + np.arange(SPP)[None, :]).ravel()
            n = ids.size
            tp, acc, alive = np.ones((3, n)), np.zeros((3, n)), np.ones(n, bool)
            tie_any = np.zeros(n, bool)
        rays = np.ascontiguousarray(rays_all[:, ids])
        gh = {k: v[..., ids] for k, v in gh_all.items()}
        live = alive.copy()
        sub = np.flatnonzero(live)
        oh = {"elem": np.full(n, -1, np.int32), "face": np.full(n, -1, np.int32), "t": np.zeros(n),
              "rst": np.zeros((3, n))}
        if sub.size:
            o_sub, _ = orc.trace(np.ascontiguousarray(rays[:, sub]), [])
            for k in oh:
                oh[k][..., sub] = o_sub[k]
            tie = np.zeros(n, bool)
            tie[sub] = parity.geometry_ties(o_sub)
            tie_any |= tie
            rep["traced"] += int(sub.size)
        ok = live & ~tie_any
        assert np.array_equal(gh["elem"][ok] >= 0, oh["elem"][ok] >= 0), depth
        both = ok & (oh["elem"] >= 0)
        assert np.array_equal(gh["elem"][both], oh["elem"][both]), depth
        assert np.array_equal(gh["face"][both], oh["face"][both]), depth
        if both.any():
            trel = np.abs(gh["t"][both] - oh["t"][both]) / oh["t"][both]
            assert trel.max() < 1e-6, depth
            rep["t_rel_max"] = max(rep["t_rel_max"], float(trel.max()))
        rep["hits"] += int(both.sum())
        dties = np.zeros(n, bool)
        nxt = po.scatter(orc, rays, oh, specs, seed, depth, depth == DEPTH - 1, tp, acc, alive, dties, ids=ids)
        tie_any |= dties
        fr.scatter(mesh, mats, seed, depth, depth == DEPTH - 1)
        g_next = fr.get_rays()[:, ids]
        ok = alive & ~tie_any
        if ok.any():
            err = float(np.max(np.abs(g_next[:, ok] - nxt[:, ok])))
            assert err < 2e-4, depth
            rep["dir_err_max"] = max(rep["dir_err_max"], err)
        assert np.array_equal(g_next[:, ~alive], nxt[:, ~alive].astype(np.float32)), depth
    img = fr.resolve(W, H, SPP)
    pix = ids[::SPP] // SPP
    ref = po.resolve(acc, len(pix), 1, SPP)
    clean = ~tie_any.reshape(len(pix), SPP).any(axis=1)
    for s in (0, 8, 16):
        a = (img[pix][clean] >> s) & 0xFF
        b = (ref[clean] >> s) & 0xFF
        assert np.max(np.abs(a.astype(int) - b.astype(int)), initial=0) <= 1
    rep["ties"] = int(tie_any.sum())
    return img, rep


def test_config3_64_meshes_one_graph_full_size(femrt, oracle, gpu_ctx):
    import oracle.parity as parity
    ctx = gpu_ctx
    specs = femrt.DEFAULT_MATERIALS
    mats = femrt.materials(specs)
    cam = femrt.auto_camera([0, 0, 0], [1, 1, 1])
    cam12 = femrt.camera_basis(cam[:3], cam[3:6], cam[6:9], float(cam[9]), W / H)
    rng = np.random.default_rng(SEED + 8)
    meshes, frames, eager, sizes = [], [], [], []
    totals = {"traced": 0, "hits": 0, "ties": 0, "t_rel_max": 0.0, "dir_err_max": 0.0}
    try:
        for i, n, p, hm in _batch_meshes(femrt):
            m = femrt.Mesh(ctx, hm)
            fr = femrt.Frame(ctx, W * H * SPP)
            meshes.append(m)
            frames.append(fr)
            sizes.append((n ** 3, p))
            img, rep = _eager_with_oracle(femrt, oracle, parity, m, hm, fr, cam12, mats, specs, SEED + i, rng)
            eager.append(img)
            for k in ("traced", "hits", "ties"):
                totals[k] += rep[k]
            for k in ("t_rel_max", "dir_err_max"):
                totals[k] = max(totals[k], rep[k])
            print(json.dumps({"mesh": i, "n_hex": n ** 3, "P": p, **rep}), flush=True)
        assert min(s for s, _ in sizes) >= 512 and max(s for s, _ in sizes) <= 32768
        assert {p for _, p in sizes} <= set(range(2, 9))
        assert totals["hits"] > 4 * N_MESH     # bounces reach the mesh again
        b = femrt.Batch(ctx)
        for i in range(N_MESH):
            b.add_paths(meshes[i], frames[i], cam12, W, H, SPP, DEPTH, mats, SEED + i)
        b.capture()
        for rep in range(2):
            # eager work overwrites every frame (other seed, one bounce)
            for i in range(N_MESH):
                frames[i].render_path(meshes[i], cam12, W, H, SPP, 1, mats, SEED + 1000 + rep)
            ctx.sync()
            b.launch()
            ctx.sync()
            for i in range(N_MESH):
                got = frames[i].rgba()[:W * H]
                np.testing.assert_array_equal(got, eager[i], err_msg=f"replay {rep} mesh {i}")
            print(json.dumps({"replay": rep, "graph_equals_eager_meshes": N_MESH}), flush=True)
        b.close()
        print(json.dumps({"config3_oracle_totals": totals}), flush=True)
    finally:
        for o in frames + meshes:
            o.close()
