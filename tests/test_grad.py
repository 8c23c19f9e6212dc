"""World gradient of the field at the isosurface crossings (the iso normal
of the reference's shading, raytracer.cu:207-216): the element-binned
matrix-core engine and the per-lane vector engine against each other and
against the FP64 oracle's gradient (fem/HexElement.cu:274-330) at the
crossings' reference coordinates, for the field orders the kernels
instantiate."""
from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("q,p,waves", [(4, 6, 0), (4, 8, 8), (3, 2, 0), (3, 4, 0)])
def test_iso_gradients_mfma_and_valu_match_oracle(femrt, oracle, gpu_ctx, q, p, waves):
    ctx = gpu_ctx
    hm = femrt.gen_structured_hex(10, q, p if not waves else 2, pert=2, amp=0.08, seed=31 + p)
    if waves:
        hm = femrt.gen_field_waves(hm, p, n_waves=waves, seed=37)
    isos = [-0.2, 0.0, 0.15] if not waves else [-0.6, -0.2, 0.2, 0.6]
    mesh = femrt.Mesh(ctx, hm)
    W, H = 256, 192
    fr = femrt.Frame(ctx, W * H)
    fr.camera_rays(femrt.auto_camera([0, 0, 0], [1, 1, 1]), W, H)
    fr.trace(mesh, femrt.trace_params(isos, guess=femrt.GUESS_WARM))
    gi = fr.iso(len(isos))
    found = gi["elem"] >= 0
    assert found.sum() > 2000
    ctx.timing(True)
    ctx.timing_reset()
    gm = fr.iso_gradients(mesh, len(isos), femrt.GRAD_MFMA)
    gv = fr.iso_gradients(mesh, len(isos), femrt.GRAD_VALU)
    kt = ctx.timing_read()
    ctx.timing(False)
    # NaN exactly where there is no crossing
    assert np.array_equal(np.isnan(gm[:, 0, :]), ~found) and np.array_equal(np.isnan(gv[:, 0, :]), ~found)
    a = np.transpose(gm, (1, 0, 2))[:, found]
    b = np.transpose(gv, (1, 0, 2))[:, found]
    mag = np.linalg.norm(b, axis=0)
    rel_mv = float(np.max(np.linalg.norm(a - b, axis=0) / np.maximum(mag, 1e-30)))
    assert rel_mv < 2e-4
    # a sample against the oracle at the FP64 crossing
    orc = oracle.Oracle.from_mesh(hm)
    ks, rs = np.nonzero(found)
    sel = np.random.default_rng(p).choice(ks.size, 300, replace=False)
    worst = 0.0
    for i in sel:
        k, r = ks[i], rs[i]
        e = int(gi["elem"][k, r])
        _, _, gw = orc.field_at(e, gi["rst"][k, :, r])
        g = gm[k, :, r].astype(np.float64)
        worst = max(worst, float(np.linalg.norm(g - gw) / max(np.linalg.norm(gw), 1e-30)))
    print(json.dumps({"q": q, "p": p, "crossings": int(found.sum()), "mfma_vs_valu_rel": rel_mv,
                      "mfma_vs_oracle_rel": worst,
                      "ms": {k: v[0] for k, v in kt.items() if k.startswith("k_grad")}}), flush=True)
    assert worst < 5e-4
    fr.close()
    mesh.close()
