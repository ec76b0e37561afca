"""Batched multistart pose registration on the device.

``DeviceAdam`` keeps per-pose Adam state in HBM and steps every pose with one
launch (``fvr_adam_step``): the update of ``gradient_optimizer.py:45-83`` per
pose, optionally with the gradient clamp and cosine learning-rate factor of
the reference's multistart driver
(``experiments/exp30_gpu_multistart_registration.py:140-183``).

``AdamLoop`` runs ``steps`` x (loss + gradient of every pose, Adam step) over
a ``BiplanarEvaluator`` or a ``RaggedBiplanarEvaluator`` -- two launches per
step: the fused kernel, then ``fvr_pn_step`` (reduce + combine, Adam, the next
items) -- eagerly, or captured once into a HIP graph (``torch.cuda.CUDAGraph``)
and replayed, so that a 100-step registration of hundreds of poses is one graph
launch.

``autotune`` picks the fused kernel's launch plan (tile height, tile order,
persistence) for an evaluator by timing each candidate on the device; choices
are cached per workload key for the process.
"""

from __future__ import annotations

import math
import threading
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .sncc import BiplanarEvaluator, LaunchPlan, RaggedBiplanarEvaluator


class DeviceAdam:
    """Per-pose Adam state on the device and its one-launch step.

    Defaults are the reference's ``register_adam`` (beta 0.9/0.999, eps 1e-8,
    lr 0.25 deg / 0.35 for rotations / translations, poses clipped to +-20 deg
    and +-25, early stop per pose when max|step| < tol).
    """

    def __init__(self, n: int, device, lr_rot: float = 0.25, lr_trans: float = 0.35, beta1: float = 0.9,
                 beta2: float = 0.999, eps: float = 1e-8, tol: float = 1e-4, rot_limit: float = 20.0,
                 trans_limit: float = 25.0, gclip_rot: float = 0.0, gclip_trans: float = 0.0, cosine_t: int = 0):
        if n < 1:
            raise ValueError("empty pose batch")
        self.n = int(n)
        self.device = device
        cfg = np.zeros(1, dtype=_lib.ADAM_DTYPE)
        cfg["lr"] = [lr_rot] * 3 + [lr_trans] * 3
        cfg["gclip"] = [gclip_rot] * 3 + [gclip_trans] * 3
        cfg["lo"] = [-rot_limit] * 3 + [-trans_limit] * 3
        cfg["hi"] = [rot_limit] * 3 + [trans_limit] * 3
        cfg["beta1"], cfg["beta2"], cfg["eps"], cfg["tol"] = beta1, beta2, eps, tol
        cfg["cosine_t"] = int(cosine_t)
        self.cfg = cfg
        f64 = dict(dtype=torch.float64, device=device)
        self.m = torch.zeros((n, 6), **f64)
        self.v = torch.zeros((n, 6), **f64)
        self.t = torch.zeros(n, dtype=torch.int32, device=device)
        self.done = torch.zeros(n, dtype=torch.int32, device=device)
        self.best_loss = torch.full((n,), math.inf, **f64)
        self.best_pose = torch.zeros((n, 6), **f64)

    def reset(self, poses: torch.Tensor) -> None:
        """Zero the moments and counters; best = (inf, current pose).  Fill kernels only."""
        self.m.zero_()
        self.v.zero_()
        self.t.zero_()
        self.done.zero_()
        self.best_loss.fill_(math.inf)
        self.best_pose.copy_(poses)

    def step(self, poses: torch.Tensor, loss: torch.Tensor, grad: torch.Tensor) -> None:
        """One Adam step of every pose in place (stream-ordered, graph-capturable)."""
        if poses.dtype != torch.float64 or tuple(poses.shape) != (self.n, 6) or not poses.is_contiguous():
            raise ValueError(f"poses must be a contiguous ({self.n}, 6) float64 tensor")
        _lib.check(_lib.lib().fvr_adam_step(_lib.ptr(poses), _lib.ptr(loss), _lib.ptr(grad), _lib.ptr(self.m),
                                            _lib.ptr(self.v), _lib.ptr(self.t), _lib.ptr(self.done),
                                            _lib.ptr(self.best_loss), _lib.ptr(self.best_pose), self.n,
                                            self.cfg.ctypes.data, _lib.stream_ptr(self.device)), "fvr_adam_step")


class AdamLoop:
    """``steps`` x (evaluate, Adam step) of a pose batch, optionally as one HIP graph.

    ``run(init_poses)`` returns (best_pose (N, 6), best_loss (N,), last_pose (N, 6))
    as device tensors.  With ``graph=True`` the loop is captured on the first
    run (after one eager warm-up step that sizes every buffer) and replayed on
    later runs.  The graph holds the evaluator's buffers by address and its target
    energies by value: whenever the evaluator's ``state_key()`` has changed since
    the capture (a new launch plan or tile map, new targets, a workspace that grew
    for a larger batch), ``run`` captures the loop again instead of replaying.
    """

    def __init__(self, evaluator, n: int, steps: int, graph: bool = True, **adam_kw):
        self.ev = evaluator
        self.n = int(n)
        self.steps = int(steps)
        self.use_graph = bool(graph)
        self.device = evaluator.device
        self.poses = torch.zeros((self.n, 6), dtype=torch.float64, device=self.device)
        self.adam = DeviceAdam(self.n, self.device, **adam_kw)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.captured_key: Optional[Tuple] = None
        self.captures = 0
        self._held: List[torch.Tensor] = []  # device buffers the current graph reads or writes
        if isinstance(evaluator, BiplanarEvaluator):
            evaluator.prepare(self.n)
            evaluator.launch_map(self.n)

    def _body(self, steps: int) -> None:
        # build the first items, then per step: the fused kernel + one launch that reduces,
        # steps Adam and builds the next items (two launches per step)
        self.ev.begin(self.poses)
        for _ in range(steps):
            self.ev.step(self.poses, self.adam)

    def _buffers(self) -> List[torch.Tensor]:
        ev = self.ev
        held = list(ev.ws.bufs.values())
        if isinstance(ev, RaggedBiplanarEvaluator):
            held += [] if ev.map is None else [ev.map]
            for s in ev.segments:
                held += list(s.folded or s.targets)
        else:
            held += [m for m, _, _ in ev._maps.values() if m is not None]
            held += list(ev.folded or ev.targets)
        return held

    def capture(self) -> None:
        """Warm up one eager step, then capture the whole loop into a graph."""
        self.graph = None
        if isinstance(self.ev, BiplanarEvaluator):
            self.ev.prepare(self.n)
            self.ev.launch_map(self.n)
        self.adam.reset(self.poses)
        self._body(1)
        torch.cuda.synchronize(self.device)
        key = self.ev.state_key()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._body(self.steps)
        torch.cuda.synchronize(self.device)
        if self.ev.state_key() != key:
            raise RuntimeError("AdamLoop: the evaluator changed its buffers during capture")
        self.graph = g
        self.captured_key = key
        self._held = self._buffers()
        self.captures += 1

    def run(self, init_poses: torch.Tensor):
        _lib.require_device(init_poses, "AdamLoop.run")
        if tuple(init_poses.shape) != (self.n, 6):
            raise ValueError(f"init_poses must be ({self.n}, 6)")
        if self.use_graph and (self.graph is None or self.ev.state_key() != self.captured_key):
            self.capture()
        self.poses.copy_(init_poses.to(torch.float64))
        self.adam.reset(self.poses)
        if self.use_graph:
            self.graph.replay()
        else:
            self._body(self.steps)
        return self.adam.best_pose, self.adam.best_loss, self.poses


# ---------------------------------------------------------------------------
# Autotuner of the fused kernel's launch plan
# ---------------------------------------------------------------------------
_CACHE: Dict[Tuple, LaunchPlan] = {}
_LOCK = threading.Lock()


def candidate_plans(flags: int) -> List[LaunchPlan]:
    """Every plan the library has a kernel for: tile height x tile order x persistence x depth."""
    L = _lib.lib()
    out = []
    for th in (32, 64, 128):
        if not L.fvr_sncc_has_tile_h(flags, th):
            continue
        for order in ("xcd", "target"):
            for persist in (0, 2, 4):
                for depth in (2, 3):
                    if L.fvr_sncc_has_depth(flags, th, depth):
                        out.append(LaunchPlan(th, order, persist, depth))
    return out


def workload_key(ev) -> Tuple:
    segs = ev.segments if isinstance(ev, RaggedBiplanarEvaluator) else [ev]
    counts = ev.counts if isinstance(ev, RaggedBiplanarEvaluator) else [None]
    return (torch.cuda.get_device_name(ev.device), ev.flags,
            tuple((s.packed.shape, s.packed.code, c) for s, c in zip(segs, counts)))


def time_plan(ev, poses: torch.Tensor, plan: LaunchPlan, reps: int = 5) -> float:
    """Median fused-kernel time (ms) of one plan, with HIP events on the launch stream."""
    ev.set_plan(plan)
    ts = []
    for r in range(reps + 1):
        e = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        if isinstance(ev, RaggedBiplanarEvaluator):
            ev.evaluate_device(poses, kernel_events=e)
        else:
            ev.evaluate_device(poses, kernel_events=e, want_moments=False)
        torch.cuda.synchronize(ev.device)
        if r:
            ts.append(e[0].elapsed_time(e[1]))
    return float(np.median(ts))


def autotune(ev, poses: torch.Tensor, plans: Optional[Sequence[LaunchPlan]] = None, reps: int = 5,
             use_cache: bool = True):
    """Time every candidate plan on ``poses``, leave ``ev`` on the fastest and return
    (best plan, {plan: ms}).  Cached per (device, flags, shapes, storage, counts)."""
    key = workload_key(ev)
    with _LOCK:
        hit = _CACHE.get(key) if use_cache and plans is None else None
    if hit is not None:
        ev.set_plan(hit)
        return hit, {}
    plans = list(plans) if plans is not None else candidate_plans(ev.flags)
    timings = {p: time_plan(ev, poses, p, reps) for p in plans}
    best = min(timings, key=timings.get)
    ev.set_plan(best)
    with _LOCK:
        _CACHE[key] = best
    return best, timings
