"""femrt_eval_field_noalloc (inputs and outputs staged in the library's fixed
device array, no device allocation) against femrt_eval_field (per-call
device buffers): the same kernel on the same inputs, so the results must be
identical bit for bit, on two calls in a row (the second smaller, so stale
contents of the shared array beyond n must not leak into the results), in
FP32 and FP64; more points than the array holds are refused."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_eval_field_noalloc_matches_allocating_entry(femrt, gpu_ctx, precision):
    hm = femrt.gen_structured_hex(5, 4, 6, pert=2, amp=0.08, seed=91)
    mesh = femrt.Mesh(gpu_ctx, hm, build_bvh=False)
    rng = np.random.default_rng(17 + precision)
    for n in (femrt.NOALLOC_MAX_POINTS, 1000):
        elem = rng.integers(0, hm.n_elem, n).astype(np.int32)
        rst = rng.uniform(-1, 1, (3, n)).astype(np.float32)
        ref = mesh.eval_field(elem, rst, precision)
        got = mesh.eval_field_noalloc(elem, rst, precision)
        assert got.shape == ref.shape == (4, n)
        assert np.isfinite(ref).all()
        np.testing.assert_array_equal(got, ref)
    mesh.close()


def test_eval_field_noalloc_rejects_more_points_than_its_scratch(femrt, gpu_ctx):
    hm = femrt.gen_structured_hex(2, 2, 2, seed=3)
    mesh = femrt.Mesh(gpu_ctx, hm, build_bvh=False)
    n = femrt.NOALLOC_MAX_POINTS + 1
    with pytest.raises(femrt.FemrtError, match="65536"):
        mesh.eval_field_noalloc(np.zeros(n, np.int32), np.zeros((3, n), np.float32), 1)
    # the entry point still works after the refusal
    out = mesh.eval_field_noalloc(np.zeros(4, np.int32), np.zeros((3, 4), np.float32), 1)
    np.testing.assert_array_equal(out, mesh.eval_field(np.zeros(4, np.int32), np.zeros((3, 4), np.float32), 1))
    mesh.close()
