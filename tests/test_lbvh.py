"""BVH builder 1, the linear BVH built on the GPU (femrt_lbvh.hip), against
the host SAH build it is an alternative to (BVH.cu:9-247 is the reference's
host build; the reference has no GPU builder).

The tree's structure is checked from the root (every element in exactly one
leaf of at most 4, every box inside its parent's, depth within the
traversal stack); traces with either tree must give the same hits and
crossings, since the nearest hit does not depend on the tree, and any ray
where they differ is traced by the FP64 oracle, where both must meet the
bar (ties aside).  Build times of both builders are reported at 32^3 and
64^3 elements."""
from __future__ import annotations

import json
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_tree(nodes, prims, boxes, n_elem, leaf_max=4):
    ab = nodes.view(np.int32)
    seen = np.zeros(n_elem, int)
    depth_max = 0
    stack = [(0, 0)]
    while stack:
        i, depth = stack.pop()
        depth_max = max(depth_max, depth)
        lo, hi = nodes[i, :3], nodes[i, 4:7]
        a, b = ab[i, 3], ab[i, 7]
        if b < 0:
            assert 1 <= -b <= leaf_max
            for s in range(a, a - b):
                e = prims[s]
                seen[e] += 1
                assert np.all(boxes[e, 0, :3] >= lo) and np.all(boxes[e, 0, 3:] <= hi)
            continue
        for c in (a, b):
            assert np.all(nodes[c, :3] >= lo) and np.all(nodes[c, 4:7] <= hi)
            stack.append((c, depth + 1))
    assert np.all(seen == 1)
    return depth_max


@pytest.mark.parametrize("n,q,seed", [(1, 3, 1), (2, 3, 2), (6, 3, 3), (11, 4, 4)])
def test_lbvh_tree_invariants(femrt, gpu_ctx, n, q, seed):
    hm = femrt.gen_structured_hex(n, q, 2, pert=2, amp=0.08, seed=seed)
    m = femrt.Mesh(gpu_ctx, hm, build_bvh=False)
    m.build_bvh(femrt.BVH_GPU_LBVH)
    nodes, prims = m.bvh()
    assert m.n_nodes == max(1, 2 * hm.n_elem - 1)
    assert sorted(prims.tolist()) == list(range(hm.n_elem))
    depth = _check_tree(nodes, prims, m.boxes(), hm.n_elem)
    assert depth < 64
    # rebuilding gives the same tree (the build is deterministic)
    m.build_bvh(femrt.BVH_GPU_LBVH)
    nodes2, prims2 = m.bvh()
    np.testing.assert_array_equal(nodes.view(np.int32), nodes2.view(np.int32))
    np.testing.assert_array_equal(prims, prims2)
    with pytest.raises(femrt.FemrtError):
        m.build_bvh(7)
    m.close()


def test_lbvh_traces_equal_sah(femrt, oracle, gpu_ctx):
    import oracle.parity as parity
    hm = femrt.gen_structured_hex(16, 4, 6, pert=2, amp=0.08, seed=77)
    isos = [-0.1, 0.0, 0.1]
    m = femrt.Mesh(gpu_ctx, hm)
    W, H = 320, 240
    fr = femrt.Frame(gpu_ctx, W * H)
    fr.camera_rays(femrt.auto_camera([0, 0, 0], [1, 1, 1]), W, H)
    prm = femrt.trace_params(isos, guess=femrt.GUESS_WARM)
    fr.trace(m, prm)
    a_h, a_i = fr.hits(), fr.iso(len(isos))
    m.build_bvh(femrt.BVH_GPU_LBVH)
    fr.trace(m, prm)
    b_h, b_i = fr.hits(), fr.iso(len(isos))
    assert (a_h["elem"] >= 0).sum() > 10000
    bad = (a_h["elem"] != b_h["elem"]) | (a_h["face"] != b_h["face"])
    hit = a_h["elem"] >= 0
    bad |= hit & (np.abs(a_h["t"] - b_h["t"]) > 1e-12 * np.abs(a_h["t"]))
    for k in range(len(isos)):
        bad |= a_i["elem"][k] != b_i["elem"][k]
        f = a_i["elem"][k] >= 0
        bad |= f & (np.abs(a_i["t"][k] - b_i["t"][k]) > 1e-12 * np.abs(a_i["t"][k]))
    ids = np.flatnonzero(bad)
    print(json.dumps({"rays": W * H, "hits": int(hit.sum()), "differ": int(ids.size)}), flush=True)
    assert ids.size <= 1e-4 * W * H
    if ids.size:
        og, oi = oracle.Oracle.from_mesh(hm).trace(np.ascontiguousarray(fr.get_rays()[:, ids]), isos)
        for gh, gi in ((a_h, a_i), (b_h, b_i)):
            sub = lambda d: {k: (v[..., ids] if isinstance(v, np.ndarray) else v) for k, v in d.items()}  # noqa: E731
            rep = parity.compare(sub(gh), sub(gi), og, oi, isos)
            assert rep["ok"], rep
    fr.close()
    m.close()


@pytest.mark.parametrize("n", [32, 64])
def test_lbvh_build_time_against_host_sah(femrt, gpu_ctx, n):
    hm = femrt.gen_structured_hex(n, 4, 2, pert=3, amp=0.15, seed=n)
    m = femrt.Mesh(gpu_ctx, hm, build_bvh=False)
    m.build_bvh(femrt.BVH_GPU_LBVH)          # first call: kernels loaded
    times = {}
    for name, b in (("host_sah", femrt.BVH_HOST_SAH), ("gpu_lbvh", femrt.BVH_GPU_LBVH)):
        t0 = time.perf_counter()
        m.build_bvh(b)
        gpu_ctx.sync()
        times[name] = (time.perf_counter() - t0) * 1e3
        nodes, prims = m.bvh()
        assert _check_tree(nodes, prims, m.boxes(), hm.n_elem) < 64
    print(json.dumps({"elements": hm.n_elem, "build_ms": times}), flush=True)
    assert times["gpu_lbvh"] < times["host_sah"]
    m.close()
