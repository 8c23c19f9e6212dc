"""Bounce stage: the CPU restatement pinned by closed forms (cosine-weighted
Lambert sampling, mirror reflection, Snell refraction on a flat face), then
the GPU's jittered camera paths, scatter and resolve against it, depth by
depth, on two separated blocks so that bounced rays hit the mesh again."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.path_oracle as po


def test_counter_rng_is_uniform_and_keyed():
    u = np.array([po.path_rand(7, p, 1, 0) for p in range(4000)])
    assert 0.47 < u.mean() < 0.53 and u.min() >= 0 and u.max() < 1
    assert po.path_rand(7, 3, 1, 0) == po.path_rand(7, 3, 1, 0)
    assert po.path_rand(7, 3, 1, 0) != po.path_rand(7, 3, 1, 1)
    assert po.path_rand(7, 3, 1, 0) != po.path_rand(8, 3, 1, 0)


def test_lambert_direction_is_cosine_weighted():
    """normal + uniform unit vector: cos(theta) of the normalised direction
    has mean 2/3 and density 2 cos(theta) on the hemisphere."""
    n = np.array([0.0, 0.0, 1.0])
    cs = []
    for p in range(20000):
        d = n + po.unit_vec(po.path_rand(3, p, 1, 0), po.path_rand(3, p, 1, 1))
        d /= np.linalg.norm(d)
        cs.append(d[2])
    cs = np.array(cs)
    assert cs.min() >= -1e-12
    assert abs(cs.mean() - 2.0 / 3.0) < 0.01
    # P(cos theta < 0.5) = 0.25 for the cosine-weighted hemisphere
    assert abs((cs < 0.5).mean() - 0.25) < 0.01


def _flat_block(femrt):
    return femrt.gen_structured_hex(1, 2, 1, pert=0, seed=0)   # unit cube, flat faces


def _one_bounce(femrt, oracle, mat, o, d, seed=5):
    hm = _flat_block(femrt)
    orc = oracle.Oracle.from_mesh(hm)
    rays = np.array([[*o, *d]], np.float32).T
    hits, _ = orc.trace(rays, [])
    tp, acc, alive = np.ones((3, 1)), np.zeros((3, 1)), np.array([True])
    nxt = po.scatter(orc, rays, hits, [mat, mat], seed, 0, False, tp, acc, alive)
    return hits, nxt, alive


def test_metal_mirror_and_dielectric_snell(femrt, oracle):
    d = np.array([0.3, 0.2, 1.0])
    d /= np.linalg.norm(d)
    # ray from below onto the t = -1 face (z = 0 plane), outward normal -z
    hits, nxt, alive = _one_bounce(femrt, oracle, (1, (0.9, 0.9, 0.9), (0, 0, 0), 0.0, 0.0), (0.5, 0.5, -1.0), d)
    assert hits["face"][0] == 0 and alive[0]
    np.testing.assert_allclose(nxt[3:, 0], [d[0], d[1], -d[2]], atol=1e-12)   # mirror
    np.testing.assert_allclose(nxt[:3, 0], [0.5 + d[0] / d[2], 0.5 + d[1] / d[2], 0.0], atol=1e-7)
    ior = 1.5
    hits, nxt, alive = _one_bounce(femrt, oracle, (2, (1, 1, 1), (0, 0, 0), 0.0, ior), (0.5, 0.5, -1.0), d,
                                   seed=11)
    out = nxt[3:, 0]
    d = d.astype(np.float32).astype(np.float64)   # the traced ray is FP32
    d /= np.linalg.norm(d)
    if out[2] > 0:   # refracted (the Fresnel draw decides); Snell: sin_i = ior sin_t
        sin_i = np.linalg.norm(d[:2])
        sin_t = np.linalg.norm(out[:2])
        assert sin_i == pytest.approx(ior * sin_t, rel=1e-7)   # FP32 ray: |d| = 1 to 1e-8
    else:
        np.testing.assert_allclose(out, [d[0], d[1], -d[2]], atol=1e-12)


def _two_blocks(femrt):
    """Two 2x1x1-element curved blocks side by side with a gap of 0.5 in x."""
    hm = femrt.gen_structured_hex(2, 3, 3, pert=1, amp=0.04, seed=5)
    g = hm.geom.reshape(hm.n_elem, 3, 4, 4, 4)
    keep = [e for e in range(hm.n_elem)]
    shift = np.array([e % 2 == 1 for e in keep])   # elements with i = 1
    for e in np.flatnonzero(shift):
        hm.vertices[e, :, 0] += 0.5
        g[e, 0, :2, :2, :2] += 0.5                  # vertex modes of x carry the constant
    hm.geom = np.ascontiguousarray(g.reshape(hm.n_elem, -1))
    adj = hm.adjacency.copy()
    for e in range(hm.n_elem):
        for f in range(6):
            ne = adj[e, f, 0]
            if ne >= 0 and shift[e] != shift[ne]:
                adj[e, f] = -1
    hm.adjacency = adj
    return hm


MATERIAL_SETS = {
    # material 0 on faces 1..5, material 1 on local face 0 (raytracer.cu:248)
    "lambert_metal": [(0, (0.7, 0.6, 0.5), (0, 0, 0), 0.0, 0.0),
                      (1, (0.8, 0.8, 0.9), (0, 0, 0), 0.2, 0.0)],
    "dielectric_lambert": [(2, (1.0, 1.0, 1.0), (0, 0, 0), 0.0, 1.5),
                           (0, (0.6, 0.7, 0.5), (0, 0, 0), 0.0, 0.0)],
}


@pytest.mark.gpu
@pytest.mark.parametrize("matset", sorted(MATERIAL_SETS))
def test_paths_match_oracle_depth_by_depth(femrt, oracle, gpu_ctx, matset):
    import oracle.parity as parity
    hm = _two_blocks(femrt)
    mesh = femrt.Mesh(gpu_ctx, hm)
    orc = oracle.Oracle.from_mesh(hm)
    specs = MATERIAL_SETS[matset]
    mats = femrt.materials(specs)
    W, H, spp, depth_max, seed = 40, 30, 2, 3, 1234
    cam = femrt.camera_basis((3.2, 2.0, 2.6), (1.25, 0.5, 0.5), (0, 1, 0), 35.0, W / H)
    fr = femrt.Frame(gpu_ctx, W * H * spp)
    fr.camera_rays_jitter(cam, W, H, spp, seed)
    fr.path_begin()
    n = W * H * spp
    tp, acc, alive = np.ones((3, n)), np.zeros((3, n)), np.ones(n, bool)
    tie_any = np.zeros(n, bool)
    prm = femrt.trace_params([], guess=femrt.GUESS_WARM, iso_enable=False, flags=0)
    rehits = 0
    for depth in range(depth_max):
        rays = fr.get_rays()
        fr.trace(mesh, prm)
        gh = fr.hits()
        oh, _ = orc.trace(rays, [])
        tie = parity.geometry_ties(oh)
        tie_any |= tie
        live = alive & ~tie_any
        assert np.array_equal(gh["elem"][live] >= 0, oh["elem"][live] >= 0), depth
        both = live & (oh["elem"] >= 0)
        assert np.array_equal(gh["elem"][both], oh["elem"][both]) and np.array_equal(gh["face"][both], oh["face"][both])
        if both.any():
            assert np.max(np.abs(gh["t"][both] - oh["t"][both]) / oh["t"][both]) < 1e-6
        if depth > 0:
            rehits += int(both.sum())
        dties = np.zeros(n, bool)
        nxt = po.scatter(orc, rays, oh, specs, seed, depth, depth == depth_max - 1, tp, acc, alive,
                         dties)
        tie_any |= dties
        fr.scatter(mesh, mats, seed, depth, depth == depth_max - 1)
        g_next = fr.get_rays()
        ok = alive & ~tie_any
        if ok.any():   # (none survive the last depth)
            assert np.max(np.abs(g_next[:, ok] - nxt[:, ok])) < 2e-4, depth
        # ended paths parked alike (origin 1e30 in FP32, direction +x)
        assert np.array_equal(g_next[:, ~alive], nxt[:, ~alive].astype(np.float32))
    assert rehits > 20          # the bounces reach the other block
    img = fr.resolve(W, H, spp)
    ref = po.resolve(acc, W, H, spp)
    clean = ~tie_any.reshape(W * H, spp).any(axis=1)
    for s in (0, 8, 16):
        a = (img[clean] >> s) & 0xFF
        b = (ref[clean] >> s) & 0xFF
        assert np.max(np.abs(a.astype(int) - b.astype(int))) <= 1
    # the one-call form gives the same image
    img2 = fr.render_path(mesh, cam, W, H, spp, depth_max, mats, seed)
    np.testing.assert_array_equal(img, img2)
    fr.close()
    mesh.close()


@pytest.mark.gpu
@pytest.mark.parametrize("W,H,cam", [
    (64, 48, ((3.0, 2.0, 2.5), (0.5, 0.5, 0.5), (0, 1, 0), 20.0)),
    (31, 77, ((-1.5, 0.4, -2.0), (0.2, 0.6, 0.4), (0, 1, 0), 35.0)),
    (128, 16, ((0.5, 4.0, 0.3), (0.5, 0.5, 0.5), (1, 0, 0), 60.0)),
])
def test_gpu_camera_rays_match_oracle(femrt, oracle, gpu_ctx, W, H, cam):
    """Pixel-centre pinhole rays generated on the device (femrt_frame_camera_rays)
    against the oracle's FP64 camera (raytracer.h:31-48), for several aspect
    ratios, field-of-view angles and up vectors; ray p = y W + x, y = 0 the
    bottom row."""
    cam10 = np.array([*cam[0], *cam[1], *cam[2], cam[3]], np.float64)
    fr = femrt.Frame(gpu_ctx, W * H)
    fr.camera_rays(cam10.astype(np.float32), W, H)
    g = fr.get_rays().astype(np.float64)
    ref = oracle.camera_rays(cam10.astype(np.float32).astype(np.float64), W, H)
    np.testing.assert_allclose(g[:3], ref[:3], rtol=0, atol=1e-6)
    np.testing.assert_allclose(g[3:], ref[3:], rtol=0, atol=2e-6)
    # corner pixels point where the closed form says: the bottom-left pixel
    # centre lies half a pixel inside the lower-left corner of the image plane
    assert np.allclose(np.linalg.norm(g[3:], axis=0), 1.0, atol=1e-6)
    fr.close()
