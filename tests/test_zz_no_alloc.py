"""The allocation-free entry point fvr_pn_evaluate_noalloc against fvr_pn_evaluate.

Both run the same kernels (set-up, fused kernel, reduce + combine); the first with its
items, Jacobians and tile partials in the library's static per-GPU scratch array, the
second in a caller workspace: losses and gradients must be equal bit for bit, on two
calls in a row (the second with fewer poses, so it runs over the first call's leftover
scratch contents), for FP16 and complex64 storage.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _evaluator(storage):
    from fvr import DifferentiableFourierVolumeRenderer
    from fvr.phantom import ellipsoid_phantom
    from fvr.sncc import BiplanarEvaluator
    vol = ellipsoid_phantom(48, n_ellipsoids=6, seed=17).astype(np.float64)
    r = DifferentiableFourierVolumeRenderer(vol, storage=storage, device=DEV)
    tp = np.random.default_rng(17).uniform(-4, 4, 6)
    ta = torch.from_numpy(r.extract_frequency_slice(*tp[:3], tx=tp[3], ty=tp[5], view="ap")).to(DEV)
    tl = torch.from_numpy(r.extract_frequency_slice(*tp[:3], tx=tp[4], ty=tp[5], view="lateral")).to(DEV)
    return BiplanarEvaluator(r.packed, ta.to(torch.complex64), tl.to(torch.complex64), adjoint=True)


def _noalloc(ev, P, loss, grad):
    from fvr import _lib
    from fvr.sncc import LaunchPlan
    n = int(P.shape[0])
    pz, py, px = ev.packed.shape
    plan = LaunchPlan(ev.plan.tile_h, "xcd", 0, ev.plan.depth)
    return _lib.lib().fvr_pn_evaluate_noalloc(
        _lib.ptr(ev.packed.table), len(ev.packed.descs), ev.templates.ctypes.data, float(pz), float(py), float(px),
        ev.lat_swap, plan.variant(ev.packed.code, ev.flags), _lib.ptr(P), n, float(ev.t2[0]), float(ev.t2[1]),
        None, _lib.ptr(loss), _lib.ptr(grad), _lib.stream_ptr(DEV))


@pytest.mark.parametrize("storage", ["fp16", "f32"])
def test_noalloc_entry_point_equals_workspace_entry_point(storage):
    from fvr.phantom import random_poses
    ev = _evaluator(storage)
    for n, seed in ((37, 1), (9, 2)):
        P = torch.from_numpy(random_poses(n, 12.0, 8.0, seed)).to(DEV)
        L0, G0 = (x.cpu().numpy().copy() for x in ev.evaluate_call(P))
        loss = torch.full((n,), float("nan"), dtype=torch.float64, device=DEV)
        grad = torch.full((n, 6), float("nan"), dtype=torch.float64, device=DEV)
        assert _noalloc(ev, P, loss, grad) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(loss.cpu().numpy(), L0, err_msg=f"{storage} n={n}")
        np.testing.assert_array_equal(grad.cpu().numpy(), G0, err_msg=f"{storage} n={n}")
        assert np.all(np.isfinite(L0)) and np.all(np.isfinite(G0))


def test_noalloc_rejects_batches_beyond_the_scratch_array():
    from fvr import _lib
    ev = _evaluator("fp16")
    L = _lib.lib()
    n = 64
    while L.fvr_pn_evaluate_workspace(ev.templates.ctypes.data, n) <= (16 << 20):
        n *= 2
    P = torch.zeros((n, 6), dtype=torch.float64, device=DEV)
    loss = torch.empty(n, dtype=torch.float64, device=DEV)
    grad = torch.empty((n, 6), dtype=torch.float64, device=DEV)
    assert _noalloc(ev, P, loss, grad) == 1
    assert b"scratch" in L.fvr_last_error()
