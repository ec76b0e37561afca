"""Powell's conjugate-direction minimiser with batched line searches.

The reference registers with ``scipy.optimize.minimize(method="Powell")``
(``core_engine.py:195-213``): Brent line searches that ask for one pose at a
time, so every evaluation is one GPU launch and one host round trip.  Here
each line search evaluates a grid of points along the direction in one call of
``loss_batch`` (one fused launch for the whole grid), zooms in around the best
grid point a fixed number of times, and finishes with a parabola through the
best point and its two neighbours (checked by one more evaluation).  The
direction set is updated as in Powell's method (the direction of largest
decrease is replaced by the sweep's net displacement when that pays, the
standard safeguard), and the stopping rule is scipy's: a sweep that lowers f by
less than ``ftol`` relative (2 (f0 - f) <= ftol (|f0| + |f|) + 1e-20) ends the
run.  Bounds are honoured exactly: every grid point is clipped to the feasible
interval of the line.
"""

from __future__ import annotations

from typing import Callable, Dict, Sequence, Tuple

import numpy as np


def _feasible(x: np.ndarray, d: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> Tuple[float, float]:
    """[a_min, a_max] such that lo <= x + a d <= hi."""
    a_lo, a_hi = -np.inf, np.inf
    for xi, di, l, h in zip(x, d, lo, hi):
        if di > 0:
            a_lo, a_hi = max(a_lo, (l - xi) / di), min(a_hi, (h - xi) / di)
        elif di < 0:
            a_lo, a_hi = max(a_lo, (h - xi) / di), min(a_hi, (l - xi) / di)
    return a_lo, a_hi


def line_search_batched(loss_batch: Callable[[np.ndarray], np.ndarray], x: np.ndarray, fx: float, d: np.ndarray,
                        lo: np.ndarray, hi: np.ndarray, step: float, n_grid: int = 16, zoom: int = 3,
                        expand: int = 3) -> Tuple[float, float, int]:
    """Minimise f(x + a d) over the feasible interval: grid rounds of ``n_grid`` points
    (one batched call each), expanding the grid while the best point is on its edge and
    the feasible interval allows, then ``zoom`` rounds around the best point, then a
    parabolic step.  Returns (a, f(x + a d), evaluations); a = 0 when nothing is lower."""
    a_lo, a_hi = _feasible(x, d, lo, hi)
    if not (a_hi > a_lo):
        return 0.0, fx, 0
    best_a, best_f, n_eval = 0.0, fx, 0
    center, half = 0.0, step
    rounds = 0
    while True:
        grid = np.clip(center + np.linspace(-half, half, n_grid), a_lo, a_hi)
        grid = np.unique(grid)
        f = np.asarray(loss_batch(x[None, :] + grid[:, None] * d[None, :]), dtype=np.float64)
        n_eval += len(grid)
        k = int(np.argmin(f))
        if f[k] < best_f:
            best_a, best_f = float(grid[k]), float(f[k])
        rounds += 1
        at_edge = (k == 0 and grid[0] > a_lo) or (k == len(grid) - 1 and grid[-1] < a_hi)
        if at_edge and rounds <= expand:
            center, half = best_a, half * 4.0
            continue
        break
    spacing = 2.0 * half / (n_grid - 1)
    last = None
    for _ in range(zoom):
        center, half = best_a, spacing
        grid = np.unique(np.clip(center + np.linspace(-half, half, n_grid), a_lo, a_hi))
        f = np.asarray(loss_batch(x[None, :] + grid[:, None] * d[None, :]), dtype=np.float64)
        n_eval += len(grid)
        k = int(np.argmin(f))
        if f[k] < best_f:
            best_a, best_f = float(grid[k]), float(f[k])
        spacing = 2.0 * half / (n_grid - 1)
        if 0 < k < len(grid) - 1:
            last = (grid[k - 1:k + 2], f[k - 1:k + 2])
        else:
            last = None
    if last is not None:  # parabola through the best grid point and its neighbours
        (a0, a1, a2), (f0, f1, f2) = last
        den = (a0 - a1) * (a0 - a2) * (a1 - a2)
        if den != 0.0:
            A = (a2 * (f1 - f0) + a1 * (f0 - f2) + a0 * (f2 - f1)) / den
            B = (a2 * a2 * (f0 - f1) + a1 * a1 * (f2 - f0) + a0 * a0 * (f1 - f2)) / den
            if A > 0:
                a_p = float(np.clip(-B / (2 * A), a_lo, a_hi))
                if a0 < a_p < a2 and a_p != best_a:
                    f_p = float(np.asarray(loss_batch((x + a_p * d)[None, :]), dtype=np.float64)[0])
                    n_eval += 1
                    if f_p < best_f:
                        best_a, best_f = a_p, f_p
    return best_a, best_f, n_eval


def batched_powell(loss_batch: Callable[[np.ndarray], np.ndarray], x0: Sequence[float],
                   bounds: Sequence[Tuple[float, float]], max_iter: int = 50, ftol: float = 1e-4,
                   step: float = 2.0, n_grid: int = 16, zoom: int = 3) -> Dict:
    """Powell's method over (n,) parameters with batched line searches.

    ``loss_batch(P)`` maps a (k, n) array of points to (k,) losses (one call per grid
    round).  Returns a dict like scipy's OptimizeResult: x, fun, nit, nfev, success,
    and ``calls`` (the number of batched calls)."""
    x = np.clip(np.asarray(x0, dtype=np.float64), [b[0] for b in bounds], [b[1] for b in bounds])
    lo = np.array([b[0] for b in bounds], dtype=np.float64)
    hi = np.array([b[1] for b in bounds], dtype=np.float64)
    n = len(x)
    dirs = np.eye(n)
    calls = [0]

    def fb(P):
        calls[0] += 1
        return loss_batch(P)

    fx = float(np.asarray(fb(x[None, :]))[0])
    nfev = 1
    nit = 0
    success = False
    steps = np.full(n, float(step))
    for nit in range(1, max_iter + 1):
        x_start, f_start = x.copy(), fx
        big_i, big_drop = 0, 0.0
        for i in range(n):
            f_before = fx
            a, fx_new, ne = line_search_batched(fb, x, fx, dirs[i], lo, hi, steps[i], n_grid, zoom)
            nfev += ne
            if fx_new < fx:
                x = np.clip(x + a * dirs[i], lo, hi)
                fx = fx_new
                steps[i] = max(abs(a), 1e-3 * step)
            if f_before - fx > big_drop:
                big_i, big_drop = i, f_before - fx
        if 2.0 * (f_start - fx) <= ftol * (abs(f_start) + abs(fx)) + 1e-20:
            success = True
            break
        # Powell's direction update: try the sweep's net displacement as a new direction
        disp = x - x_start
        if np.any(disp != 0.0):
            x_ext = np.clip(x + disp, lo, hi)
            f_ext = float(np.asarray(fb(x_ext[None, :]))[0])
            nfev += 1
            if f_ext < f_start:
                t = 2.0 * (f_start - 2.0 * fx + f_ext) * (f_start - fx - big_drop) ** 2 - big_drop * (f_start - f_ext) ** 2
                if t < 0.0:
                    a, fx_new, ne = line_search_batched(fb, x, fx, disp, lo, hi, 1.0, n_grid, zoom)
                    nfev += ne
                    if fx_new < fx:
                        x = np.clip(x + a * disp, lo, hi)
                        fx = fx_new
                    dirs[big_i] = dirs[-1]
                    dirs[-1] = disp
                    steps[big_i] = steps[-1]
                    steps[-1] = 1.0
    return {"x": x, "fun": fx, "nit": nit, "nfev": nfev, "success": success, "calls": calls[0]}
