"""CPU tests of the batched Powell minimiser (fvr/powell.py) on closed-form objectives."""

import numpy as np
import pytest
from scipy.optimize import minimize


def _batch(f):
    calls = []

    def fb(P):
        P = np.atleast_2d(P)
        calls.append(len(P))
        return np.array([f(p) for p in P])

    return fb, calls


def test_coupled_quadratic_reaches_its_minimum():
    from fvr.powell import batched_powell
    rng = np.random.default_rng(2)
    Q = rng.standard_normal((6, 6))
    H = Q @ Q.T + 6 * np.eye(6)
    xs = rng.uniform(-3, 3, 6)
    f = lambda x: 0.5 * (x - xs) @ H @ (x - xs) + 1.0
    fb, calls = _batch(f)
    res = batched_powell(fb, np.zeros(6), [(-20, 20)] * 6, max_iter=50, ftol=1e-10)
    assert res["success"]
    assert np.abs(res["x"] - xs).max() < 1e-3, (res["x"], xs)
    assert abs(res["fun"] - 1.0) < 1e-6
    assert res["calls"] == len(calls) and sum(calls) == res["nfev"]
    assert res["nfev"] / res["calls"] > 8  # most points come in batches


def test_bounds_are_honoured_and_the_minimum_lands_on_them():
    from fvr.powell import batched_powell
    xs = np.array([5.0, -7.0, 0.5, 30.0, -40.0, 2.0])  # two coordinates outside the box
    f = lambda x: float(np.sum((x - xs) ** 2 * np.arange(1, 7)))
    fb, _ = _batch(lambda x: (f(x) if np.all(np.abs(x) <= np.array([20] * 3 + [25] * 3) + 1e-12) else np.nan))
    res = batched_powell(fb, np.zeros(6), [(-20, 20)] * 3 + [(-25, 25)] * 3, max_iter=50, ftol=1e-12)
    np.testing.assert_allclose(res["x"], np.clip(xs, [-20] * 3 + [-25] * 3, [20] * 3 + [25] * 3), atol=1e-3)


def test_narrow_rotated_valley_matches_scipy_powell():
    """A 6-D extended Rosenbrock function: the batched method ends at least as low as
    scipy's Powell with the same bounds and tolerance."""
    from fvr.powell import batched_powell

    def f(x):
        return float(sum(100 * (x[i + 1] - x[i] ** 2) ** 2 + (1 - x[i]) ** 2 for i in range(0, 6, 2)))

    fb, _ = _batch(f)
    x0 = np.array([-1.2, 1.0, -0.5, 0.8, 0.3, -0.4])
    bnds = [(-5, 5)] * 6
    res = batched_powell(fb, x0, bnds, max_iter=200, ftol=1e-10, step=0.5)
    ref = minimize(f, x0, method="Powell", bounds=bnds, options={"maxiter": 200, "ftol": 1e-10})
    assert res["fun"] <= max(ref.fun, 1e-8) * 10 + 1e-10, (res["fun"], ref.fun)
    assert np.abs(res["x"] - 1.0).max() < 1e-2


def test_line_search_parabola_and_feasible_interval():
    from fvr.powell import _feasible, line_search_batched
    a_lo, a_hi = _feasible(np.array([0.0, 1.0]), np.array([1.0, -2.0]), np.array([-1.0, -3.0]), np.array([2.0, 3.0]))
    assert (a_lo, a_hi) == (-1.0, 2.0)
    fb, calls = _batch(lambda x: (x[0] - 0.3137) ** 2)
    a, fa, n = line_search_batched(fb, np.zeros(2), 0.3137 ** 2, np.array([1.0, 0.0]), np.array([-5.0, -5.0]),
                                   np.array([5.0, 5.0]), 1.0)
    assert abs(a - 0.3137) < 1e-9 and fa < 1e-17  # exact on a parabola
    assert n == sum(calls)
