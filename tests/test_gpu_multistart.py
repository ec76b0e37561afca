"""GPU tests of the batched multistart machinery: launch plans of the fused kernel,
several volumes in one launch, the device Adam step, the graph-captured loop
and the autotuner.

Tolerances: for one tile height, changing the tile order or the persistence
changes only which workgroup computes a tile, so results are bit-identical;
another tile height regroups the float32 tile sums (loss within 1e-5 relative,
gradient within 1e-4 of max|grad|).  A volume inside a ragged launch computes
exactly the tiles it computes alone: bit-identical.  The Adam kernel against
the FP64 oracle on the same loss / gradient sequence: 1e-12.
"""

import numpy as np
import pytest
import torch

from oracle import diff_fvr_np as O

pytestmark = pytest.mark.gpu


def _renderer(n, seed, storage="fp16"):
    from fvr import DifferentiableFourierVolumeRenderer
    from fvr.phantom import ellipsoid_phantom
    vol = ellipsoid_phantom(n, n_ellipsoids=6, seed=seed).astype(np.float64)
    return DifferentiableFourierVolumeRenderer(vol, storage=storage, device=torch.device("cuda", 0))


def _targets(r, tp):
    a = torch.from_numpy(r.extract_frequency_slice(*tp[:3], tx=tp[3], ty=tp[5], view="ap")).to("cuda")
    b = torch.from_numpy(r.extract_frequency_slice(*tp[:3], tx=tp[4], ty=tp[5], view="lateral")).to("cuda")
    return a.to(torch.complex64), b.to(torch.complex64)


@pytest.fixture(scope="module")
def vols():
    out = []
    for n, seed in ((32, 11), (48, 12), (40, 13)):
        r = _renderer(n, seed)
        tp = np.random.default_rng(seed).uniform(-4, 4, 6)
        out.append((r, _targets(r, tp)))
    return out


def _poses(n, seed):
    from fvr.phantom import random_poses
    return torch.from_numpy(random_poses(n, 12.0, 8.0, seed)).to("cuda")


@pytest.mark.parametrize("herm", [True, False], ids=["fold", "full"])
def test_launch_plans_agree(vols, herm):
    from fvr.multistart import candidate_plans
    from fvr.sncc import BiplanarEvaluator, LaunchPlan
    r, (ta, tl) = vols[1]
    ev = BiplanarEvaluator(r.packed, ta, tl, adjoint=True, herm=herm)
    P = _poses(19, 3)
    ref = {}
    for th in (32, 64, 128):
        if not __import__("fvr")._lib.lib().fvr_sncc_has_tile_h(ev.flags, th):
            continue
        ev.set_plan(LaunchPlan(th))
        L, G = ev.evaluate_device(P)
        ref[th] = (L.cpu().numpy().copy(), G.cpu().numpy().copy())
    assert set(ref) == {32, 64, 128}
    for th in (32, 128):
        np.testing.assert_allclose(ref[th][0], ref[64][0], rtol=1e-5)
        assert np.max(np.abs(ref[th][1] - ref[64][1])) <= 1e-4 * np.max(np.abs(ref[64][1]))
    plans = candidate_plans(ev.flags)
    assert len(plans) == (36 if herm else 18)  # the fold kernel also has depth 3
    for plan in plans:
        ev.set_plan(plan)
        L, G = ev.evaluate_device(P)
        np.testing.assert_array_equal(L.cpu().numpy(), ref[plan.tile_h][0], err_msg=str(plan))
        np.testing.assert_array_equal(G.cpu().numpy(), ref[plan.tile_h][1], err_msg=str(plan))


def _fold_wave_classes(items: np.ndarray, tile_h: int) -> np.ndarray:
    """Counts of (per-sample tests, interior, outside) waves of fold items: the classes
    the fold kernel gives each wave, from the items' 32.32 geometry (used here to show that
    the poses of a test reach every class).  A wave step is a 2 x 32 patch: lane (rl, c)
    of the wave whose first row is h/2 + r0 walks rows h/2 + r0 + rl + 2t, t < R; only
    waves whose every step is a regular row (not the first wave, no row past h - 1) are
    classified."""
    R = tile_h // 4
    counts = np.zeros(3, dtype=np.int64)
    for it in items:
        h, w = int(it["h"]), int(it["w"])
        hh = h // 2
        c = np.array(it["c"], dtype=np.int64)
        u = np.array(it["ustep"], dtype=np.int64)
        v = np.array(it["vstep"], dtype=np.int64)
        hi = np.array(it["box_hi"], dtype=np.int64)
        K = hi + (1 << 32) + (hi & 0xFFFFFFFF)
        M = K - hi
        for r0 in range(0, hh, 2 * R):
            full = r0 > 0 and hh - r0 >= 2 * R
            for tx in range(int(it["tiles_x"])):
                if not full:
                    counts[0] += 1
                    continue
                j = tx * 32 + np.arange(64) % 32
                rl = np.arange(64) // 32
                colok = j < w
                jc = np.minimum(j, w - 1)
                q = c[None, :] + jc[:, None] * u[None, :] + (hh + r0 + rl)[:, None] * v[None, :]
                last = q + 2 * (R - 1) * v[None, :]
                inside = np.all(colok) and np.all((q >= M) & (q <= hi) & (last >= M) & (last <= hi))
                away = ((q < 0) & (last < 0)) | ((q > K) & (last > K))
                outside = np.all(~colok | np.any(away, axis=1))
                counts[1 if inside else (2 if outside else 0)] += 1
    return counts


@pytest.mark.parametrize("storage", ["fp16", "bf16", "f32"])
def test_fold_wave_classes_equal_per_sample_tests(storage):
    """The fold kernel's interior / outside wave classes give the same sums, bit for bit,
    as per-sample box tests on every wave, for every tile height and depth; the poses
    (rotations up to 40 degrees on a 176 x 176 slice: its last 32-column tile holds 16
    columns) reach all three classes and lanes past the last column."""
    from fvr.multistart import candidate_plans
    from fvr.sncc import BiplanarEvaluator, LaunchPlan
    r = _renderer(88, 21, storage)
    ta, tl = _targets(r, np.random.default_rng(21).uniform(-4, 4, 6))
    ev = BiplanarEvaluator(r.packed, ta, tl, adjoint=True)
    from fvr.phantom import random_poses
    P = torch.from_numpy(np.concatenate([random_poses(9, 40.0, 8.0, 5), random_poses(6, 5.0, 8.0, 6)])).to("cuda")
    seen = np.zeros(3, dtype=np.int64)
    plans = [p for p in candidate_plans(ev.flags) if p.order == "xcd" and p.persist == 0]
    assert len(plans) == 6
    for plan in plans:
        ev.set_plan(LaunchPlan(plan.tile_h, depth=plan.depth, exact_box=True))
        L0, G0 = (x.cpu().numpy().copy() for x in ev.evaluate_device(P))
        ev.set_plan(plan)
        L1, G1 = (x.cpu().numpy().copy() for x in ev.evaluate_device(P))
        np.testing.assert_array_equal(L1, L0, err_msg=str(plan))
        np.testing.assert_array_equal(G1, G0, err_msg=str(plan))
        assert np.all(np.isfinite(L1)) and np.all(np.isfinite(G1))
        seen += _fold_wave_classes(ev.device_items(len(P)), plan.tile_h)
    assert np.all(seen > 0), seen
    assert r.packed.shape[2] % 32 != 0 or r.packed.shape[1] % 32 != 0


@pytest.mark.parametrize("order", ["xcd", "target"])
def test_ragged_launch_equals_separate_launches(vols, order):
    from fvr.sncc import BiplanarEvaluator, LaunchPlan, RaggedBiplanarEvaluator
    plan = LaunchPlan(64, order, 2 if order == "target" else 0)
    counts = (5, 7, 3)
    segs = [BiplanarEvaluator(r.packed, ta, tl, adjoint=True, plan=plan) for r, (ta, tl) in vols]
    rag = RaggedBiplanarEvaluator(segs, counts, plan)
    P = _poses(sum(counts), 7)
    L, G = rag.evaluate_device(P)
    L, G = L.cpu().numpy().copy(), G.cpu().numpy().copy()
    off = 0
    for s, c in zip(segs, counts):
        l1, g1 = s.evaluate_device(P[off: off + c].contiguous())
        np.testing.assert_array_equal(L[off: off + c], l1.cpu().numpy())
        np.testing.assert_array_equal(G[off: off + c], g1.cpu().numpy())
        off += c
    # the device-built items of the ragged launch pass the host validator against the combined table
    from fvr import _items
    items = rag.device_items()
    assert _items.validate_items(items, len(rag.descs), rag.descs, rag.rows) == rag.n_tiles


def _adam_sequences(n, steps, seed):
    rng = np.random.default_rng(seed)
    losses = rng.uniform(0.1, 1.0, (steps, n))
    losses[3] = 0.01  # a new best at step 4
    grads = rng.standard_normal((steps, n, 6)) * np.array([0.3] * 3 + [3.0] * 3)
    grads[:, 0] *= 1e-12  # pose 0: |grad| << eps, so its first step is below tol (early stop)
    return losses, grads


@pytest.mark.parametrize("opts", ["reference", "exp30"])
def test_device_adam_equals_oracle(opts):
    from fvr.multistart import DeviceAdam
    n, steps = 5, 25
    losses, grads = _adam_sequences(n, steps, 3 if opts == "reference" else 4)
    p0 = np.random.default_rng(9).uniform(-15, 15, (n, 6))
    if opts == "reference":
        kw = dict(tol=1e-4)
        okw = dict(tol=1e-4)
    else:
        kw = dict(lr_rot=0.02, lr_trans=2.0, beta1=0.8, beta2=0.99, tol=0.0, rot_limit=1e9, trans_limit=1e9,
                  gclip_rot=0.25, gclip_trans=2.5, cosine_t=20)
        okw = dict(lr_rot=0.02, lr_trans=2.0, beta1=0.8, beta2=0.99, tol=0.0, limits=None, gclip=(0.25, 2.5),
                   cosine_t=20)
    dev = torch.device("cuda", 0)
    adam = DeviceAdam(n, dev, **kw)
    P = torch.from_numpy(p0.copy()).to(dev)
    adam.reset(P)
    for s in range(steps):
        adam.step(P, torch.from_numpy(losses[s]).to(dev), torch.from_numpy(grads[s]).to(dev))
    torch.cuda.synchronize()
    for k in range(n):
        it = iter(range(steps))

        def f(p, k=k, it=it):
            s = next(it)
            return float(losses[s, k]), grads[s, k]

        res = O.adam_register(f, p0[k], max_iter=steps, **okw)
        iters = res["iterations"]
        assert int(adam.t[k]) == iters
        np.testing.assert_allclose(adam.best_pose[k].cpu().numpy(), res["final_pose"], rtol=1e-12, atol=1e-12)
        assert float(adam.best_loss[k]) == res["final_loss"]
        if iters < steps:  # stopped early: the device pose froze after the same step
            assert int(adam.done[k]) == 1
        else:
            assert int(adam.done[k]) == 0
    assert int(adam.done[0]) == (1 if opts == "reference" else 0)


def test_graph_loop_equals_eager_on_every_replay(vols):
    from fvr.multistart import AdamLoop
    from fvr.sncc import BiplanarEvaluator
    r, (ta, tl) = vols[1]
    ev_g = BiplanarEvaluator(r.packed, ta, tl)
    ev_e = BiplanarEvaluator(r.packed, ta, tl)
    P0 = _poses(8, 21)
    eager = AdamLoop(ev_e, 8, 12, graph=False, tol=0.0)
    bp, bl, last = (t.cpu().numpy().copy() for t in eager.run(P0))
    loop = AdamLoop(ev_g, 8, 12, graph=True, tol=0.0)
    for rep in range(3):
        gp, gl, glast = (t.cpu().numpy().copy() for t in loop.run(P0))
        np.testing.assert_array_equal(gp, bp, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(gl, bl, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(glast, last, err_msg=f"replay {rep}")
        ev_e.evaluate_device(_poses(3, rep))  # unrelated eager work between replays
        torch.cuda.synchronize()
    assert loop.graph is not None
    assert np.all(bl < 1.0)


def test_ragged_graph_loop_and_autotune(vols):
    from fvr.multistart import AdamLoop, autotune, candidate_plans
    from fvr.sncc import BiplanarEvaluator, RaggedBiplanarEvaluator
    counts = (6, 9, 4)
    segs = [BiplanarEvaluator(r.packed, ta, tl) for r, (ta, tl) in vols]
    rag = RaggedBiplanarEvaluator(segs, counts)
    P0 = _poses(sum(counts), 5)
    best, timings = autotune(rag, P0, reps=2, use_cache=False)
    assert best in candidate_plans(rag.flags) and len(timings) == 36 and rag.plan == best
    assert all(t > 0 for t in timings.values())
    again, t2 = autotune(rag, P0)  # cached
    assert again == best and t2 == {}
    eager = AdamLoop(rag, sum(counts), 6, graph=False, tol=0.0)
    bp, bl, _ = (t.cpu().numpy().copy() for t in eager.run(P0))
    loop = AdamLoop(rag, sum(counts), 6, graph=True, tol=0.0)
    for rep in range(2):
        gp, gl, _ = (t.cpu().numpy().copy() for t in loop.run(P0))
        np.testing.assert_array_equal(gp, bp)
        np.testing.assert_array_equal(gl, bl)


@pytest.mark.parametrize("ragged", [False, True], ids=["single", "ragged"])
def test_two_launch_step_equals_separate_launches(vols, ragged):
    """begin() + step(adam) (fused kernel + fvr_pn_step) reproduces, bit for bit, the
    separate launches: set-up, fused kernel, reduce/combine, fvr_adam_step."""
    from fvr.multistart import DeviceAdam
    from fvr.sncc import BiplanarEvaluator, RaggedBiplanarEvaluator
    def make():  # one evaluator (and workspace) per path
        if ragged:
            return RaggedBiplanarEvaluator([BiplanarEvaluator(r.packed, ta, tl) for r, (ta, tl) in vols], (4, 3, 5))
        r, (ta, tl) = vols[1]
        return BiplanarEvaluator(r.packed, ta, tl)

    ev, ev2 = make(), make()
    n = 12 if ragged else 9
    P0 = _poses(n, 31)
    dev = P0.device
    pa, pb = P0.clone(), P0.clone()
    aa, ab = DeviceAdam(n, dev, tol=0.0), DeviceAdam(n, dev, tol=0.0)
    aa.reset(pa)
    ab.reset(pb)
    ev.begin(pa)
    for k in range(6):
        la, ga = (t.cpu().numpy().copy() for t in ev.step(pa, aa))
        lb, gb = ev2.evaluate_device(pb) if ragged else ev2.evaluate_device(pb, want_moments=False)
        lb, gb = lb.cpu().numpy().copy(), gb.cpu().numpy().copy()
        ab.step(pb, torch.from_numpy(lb).to(dev), torch.from_numpy(gb).to(dev))
        np.testing.assert_array_equal(la, lb, err_msg=f"step {k}")
        np.testing.assert_array_equal(ga, gb, err_msg=f"step {k}")
        np.testing.assert_array_equal(pa.cpu().numpy(), pb.cpu().numpy(), err_msg=f"step {k}")
    for x, y in ((aa.best_pose, ab.best_pose), (aa.best_loss, ab.best_loss), (aa.m, ab.m), (aa.v, ab.v),
                 (aa.t, ab.t)):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())


def test_graph_loop_recaptures_after_plan_targets_or_workspace_change(vols):
    """A captured loop is replayed only while the evaluator's plan, targets and workspace are
    the ones it was captured over: a new plan, new targets or a larger batch that regrows the
    workspace make the next run capture again, and every run equals the eager loop."""
    from fvr.multistart import AdamLoop
    from fvr.sncc import BiplanarEvaluator, LaunchPlan
    r, (ta, tl) = vols[1]
    ev = BiplanarEvaluator(r.packed, ta, tl)
    P0 = _poses(8, 23)

    def eager_ref():
        e = BiplanarEvaluator(r.packed, *(t.clone() for t in ev.targets), plan=ev.plan)
        return [t.cpu().numpy().copy() for t in AdamLoop(e, 8, 7, graph=False, tol=0.0).run(P0)]

    loop = AdamLoop(ev, 8, 7, graph=True, tol=0.0)

    def check(expect_captures):
        got = [t.cpu().numpy().copy() for t in loop.run(P0)]
        for a, b in zip(got, eager_ref()):
            np.testing.assert_array_equal(a, b)
        assert loop.captures == expect_captures, (loop.captures, expect_captures)

    check(1)
    check(1)  # plain replay
    ev.set_plan(ev.plan)  # the same plan: nothing changes
    check(1)
    ev.set_plan(LaunchPlan(32, "target", 0, 2))
    check(2)
    ev.set_targets(tl.flip(0).contiguous() if ta.shape == tl.shape else ta, tl)
    check(3)
    ev.evaluate_device(_poses(40, 3))  # regrows the workspace the graph was captured over
    check(4)
    check(4)


@pytest.mark.parametrize("storage", ["fp16", "f32"])
def test_one_call_entry_point_equals_three_launches(storage):
    """fvr_pn_evaluate (one C call: set-up, fused kernel, reduce + combine in a caller
    workspace) gives the three-launch evaluation's losses and gradients bit for bit, for
    every tile height and depth, in the default order and in the kernel's map-free
    target-major order (against the host-built target map), and a second call with fewer
    poses reuses the workspace."""
    from fvr.multistart import candidate_plans
    from fvr.sncc import BiplanarEvaluator
    r = _renderer(40, 13, storage)
    ta, tl = _targets(r, np.random.default_rng(13).uniform(-4, 4, 6))
    ev = BiplanarEvaluator(r.packed, ta, tl, adjoint=True)
    for plan in [p for p in candidate_plans(ev.flags) if p.persist == 0]:
        ev.set_plan(plan)
        for n, seed in ((11, 3), (5, 4)):
            P = _poses(n, seed)
            L0, G0 = (x.cpu().numpy().copy() for x in ev.evaluate_device(P))
            L1, G1 = (x.cpu().numpy().copy() for x in ev.evaluate_call(P))
            np.testing.assert_array_equal(L1, L0, err_msg=str(plan))
            np.testing.assert_array_equal(G1, G0, err_msg=str(plan))


def test_autotune_entry_point_picks_a_runnable_plan_and_caches_it():
    """fvr_autotune_pn_evaluate times the (tile height, depth, order) variants of
    fvr_pn_evaluate in the library and returns one the evaluator runs: the C call with it
    gives the three-launch result of that plan bit for bit; a second call is a cache hit
    (the same variant and time); after fvr_autotune_clear the search runs again and still
    returns a valid plan."""
    from fvr import _lib
    from fvr.sncc import BiplanarEvaluator
    r = _renderer(40, 13, "fp16")
    ta, tl = _targets(r, np.random.default_rng(14).uniform(-4, 4, 6))
    ev = BiplanarEvaluator(r.packed, ta, tl, adjoint=True)
    P = _poses(9, 5)
    _lib.check(_lib.lib().fvr_autotune_clear(), "fvr_autotune_clear")
    plan, ms = ev.autotune_call(P, reps=2)
    assert ms > 0 and plan.tile_h in (32, 64, 128) and plan.order in ("xcd", "target") and ev.plan == plan
    L0, G0 = (x.cpu().numpy().copy() for x in ev.evaluate_device(P))
    L1, G1 = (x.cpu().numpy().copy() for x in ev.evaluate_call(P))
    np.testing.assert_array_equal(L1, L0)
    np.testing.assert_array_equal(G1, G0)
    assert ev.autotune_call(P, reps=2) == (plan, ms)
    _lib.check(_lib.lib().fvr_autotune_clear(), "fvr_autotune_clear")
    plan3, ms3 = ev.autotune_call(P, reps=2)
    assert ms3 > 0 and plan3.tile_h in (32, 64, 128)


def test_non_finite_poses_give_nan_and_leave_the_other_poses_unchanged():
    """A pose with a NaN or infinite entry (any of the six, tz included: it enters both views'
    phase ramps) gets a NaN loss and gradient -- its items read nothing and its Jacobians
    are NaN -- through the three-launch and the one-call paths, and every other pose of the
    batch keeps its result bit for bit."""
    from fvr.sncc import BiplanarEvaluator
    r = _renderer(40, 13, "fp16")
    ta, tl = _targets(r, np.random.default_rng(15).uniform(-4, 4, 6))
    ev = BiplanarEvaluator(r.packed, ta, tl, adjoint=True)
    P = _poses(7, 21)
    L0, G0 = (x.cpu().numpy().copy() for x in ev.evaluate_device(P))
    Q = P.clone()
    Q[1, 0] = float("nan")
    Q[3, 4] = float("inf")
    Q[5, 5] = -float("inf")
    bad = [1, 3, 5]
    good = [0, 2, 4, 6]
    for run in (ev.evaluate_device, ev.evaluate_call):
        L1, G1 = (x.cpu().numpy().copy() for x in run(Q))
        assert np.all(np.isnan(L1[bad])) and np.all(np.isnan(G1[bad]))
        np.testing.assert_array_equal(L1[good], L0[good])
        np.testing.assert_array_equal(G1[good], G0[good])
