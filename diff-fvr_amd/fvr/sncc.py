"""Launchers for the fused Spectral-NCC forward + pose-adjoint kernel.

``run_items``        evaluate arbitrary host-built items (any view, optional
                     slice output) and return their float64 moment sums.
``BiplanarEvaluator`` the batched hot path of convention N: device pose
                     set-up -> fused kernel -> fixed-order reduction -> biplanar
                     combine, stream-ordered, no host sync and no allocation
                     once its workspace is sized (graph-capturable).

Moment order (per item): A, E, Mc_im, Mr_im, Mc_g[3], Mr_g[3], Mc_h[3], Mr_h[3]
(see ``csrc/fvr_sncc.hip``).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _items, _lib
from .spectrum import PackedSpectrum

TWO_PI = 2.0 * math.pi


def tile_rows() -> int:
    """Default tile height (rows) of the fused kernel."""
    return int(_lib.lib().fvr_sncc_tile_rows())


@dataclass(frozen=True)
class LaunchPlan:
    """How a batch is laid out on the fused kernel.

    tile_h: rows per tile (32, 64, 128; 4 waves x 8 / 16 / 32 rows, 64 columns); a
    Hermitian-fold tile is 32 columns x 2 tile_h rows (each wave step a 2 x 32 patch).
    order: "xcd" -- each XCD takes a contiguous range of tiles in item order
    (pose-major); "target" -- tiles sorted by (view, tile row, tile column, pose)
    before the contiguous XCD split, so that an XCD's L2 holds one band of one
    view's target while it walks every pose.
    persist: workgroups per CU (0: one workgroup per tile slot; k: k * 256
    persistent workgroups walking the slots).
    depth: samples whose loads are in flight per lane (2, or 3 for the fold kernel).
    exact_box: fold kernel only -- every wave takes the per-sample box tests instead of
    the interior / outside wave classes (same sums bit for bit; for tests).
    """
    tile_h: int = 64
    order: str = "xcd"
    persist: int = 0
    depth: int = 2  # samples in flight per lane (the Hermitian-fold kernel also has 3)
    exact_box: bool = False

    def variant(self, code: int, flags: int) -> int:
        return code | (flags << 4) | (self.tile_h << 8) | (self.depth << 16) | (int(self.exact_box) << 20)


# variant bit of the fused kernel's map-free target-major order for (AP, LAT) item pairs
VARIANT_BIPLANAR_ORDER = 1 << 21


N_CU = 256  # MI355X compute units (8 XCDs x 32)


def slot_map(entries: np.ndarray) -> np.ndarray:
    """(n_slots, 2) int32 map: slot w (XCD w % 8) takes entry (w % 8) * ceil(n/8) + w // 8
    of ``entries`` ((n, 2) = (item, tile within the item)); the padding slots are idle (-1)."""
    n = len(entries)
    per = (n + 7) // 8
    w = np.arange(8 * per)
    seq = (w % 8) * per + w // 8
    out = np.full((8 * per, 2), -1, dtype=np.int32)
    ok = seq < n
    out[ok] = entries[seq[ok]]
    return out


def target_major_entries(n_poses: int, tiles: Sequence[Tuple[int, int]], item0: int = 0) -> np.ndarray:
    """(item, tile) pairs of a biplanar segment in (view, tile row, tile column, pose) order;
    ``tiles[v]`` = (tiles_x, tiles_y) of view v, items 2 pose + view from ``item0``."""
    out = []
    for v, (tx, ty) in enumerate(tiles):
        lt = np.arange(tx * ty)
        p = np.arange(n_poses)
        item = item0 + 2 * p[None, :] + v
        out.append(np.stack(np.broadcast_arrays(item, lt[:, None]), axis=-1).reshape(-1, 2))
    return np.concatenate(out).astype(np.int32)


def item_major_entries(n_poses: int, tiles: Sequence[Tuple[int, int]], item0: int = 0) -> np.ndarray:
    out = []
    for p in range(n_poses):
        for v, (tx, ty) in enumerate(tiles):
            lt = np.arange(tx * ty)
            out.append(np.stack([np.full_like(lt, item0 + 2 * p + v), lt], axis=-1))
    return np.concatenate(out).astype(np.int32)


class _Workspace:
    """Grow-only device buffers keyed by name (no allocation once sized).  ``version``
    counts (re)allocations: a HIP graph captured over these buffers is valid only while
    it is unchanged."""

    def __init__(self, device):
        self.device = device
        self.bufs: Dict[str, torch.Tensor] = {}
        self.version = 0

    def get(self, name: str, numel: int, dtype) -> torch.Tensor:
        t = self.bufs.get(name)
        if t is None or t.numel() < numel or t.dtype != dtype:
            t = torch.empty(max(numel, 1), dtype=dtype, device=self.device)
            self.bufs[name] = t
            self.version += 1
        return t


def item_tiles(items: np.ndarray, rows: int) -> np.ndarray:
    """Tile count per item (``_items.tiles_for``: 64 x rows tiles; 32 x 2 rows for the fold)."""
    return np.array([int(np.prod(_items.tiles_for(int(it["h"]), int(it["w"]), rows,
                                                  bool(it["flags"] & _lib.ITEM_HERM)))) for it in items],
                    dtype=np.int64)


def fold_target(t: torch.Tensor) -> torch.Tensor:
    """(h, w, 4) float32 (Re T, Im T, Re, Im of conj T(h-i, w-j)) for the Hermitian-fold
    kernel; the mirror part is 0 in row 0 and column 0 (they have no mirror in the grid).
    Row h/2 carries the fold's masks: columns 1 .. w/2 - 1 are 0 (they are the mirrors of
    the right half of the same row) and column w/2 has no mirror part (it is its own)."""
    h, w = t.shape
    tc = t.to(torch.complex64)
    m = torch.zeros_like(tc)
    m[1:, 1:] = torch.conj(torch.flip(tc[1:, 1:], dims=(0, 1)))
    tc = tc.clone()
    hh, hw = h // 2, w // 2
    tc[hh, 1:hw] = 0
    m[hh, 1:hw + 1] = 0
    return torch.stack([tc.real, tc.imag, m.real, m.imag], dim=-1).contiguous()


def run_items(packed: PackedSpectrum, items: np.ndarray, flags: int,
              ws: Optional[_Workspace] = None) -> torch.Tensor:
    """Evaluate host-built items; returns (n_items, 16) float64 moments on the device.

    ``items['tile_base']`` is (re)assigned here.  Every item must carry the
    same ``flags`` (one kernel variant per launch).
    """
    items = _lib.as_struct(items, _lib.ITEM_DTYPE).copy()
    rows = int(items["tile_h"][0]) if len(items) else tile_rows()
    ntile = item_tiles(items, rows)
    items["tile_base"] = np.concatenate([[0], np.cumsum(ntile)[:-1]]).astype(np.int32)
    if np.any(items["flags"] != flags):
        raise ValueError("run_items: every item of one launch must carry the launch flags")
    n_tiles = _items.validate_items(items, len(packed.descs), packed.descs, rows)
    ws = ws or _Workspace(packed.device)
    red = torch.empty((len(items), _lib.NMOM), dtype=torch.float64, device=packed.device)
    launch_items(packed, items, n_tiles, flags, ws, red)
    return red


def launch_items(packed: PackedSpectrum, items: np.ndarray, n_tiles: int, flags: int, ws: _Workspace,
                 red: torch.Tensor) -> None:
    """Upload validated items and launch the fused kernel + fixed-order reduction (stream-ordered)."""
    L = _lib.lib()
    n_items = len(items)
    dev_items = ws.get("items", n_items * _lib.ITEM_DTYPE.itemsize, torch.uint8)
    dev_items[: n_items * _lib.ITEM_DTYPE.itemsize].copy_(torch.from_numpy(_lib.struct_bytes(items, _lib.ITEM_DTYPE)))
    partials = ws.get("partials", n_tiles * _lib.NMOM, torch.float32)
    st = _lib.stream_ptr(packed.device)
    variant = packed.code | (flags << 4) | (int(items["tile_h"][0]) << 8)
    _lib.check(L.fvr_sncc_eval_v(_lib.ptr(packed.table), len(packed.descs), _lib.ptr(dev_items), n_items,
                                 n_tiles, _lib.ptr(partials), variant, st), "fvr_sncc_eval")
    _lib.check(L.fvr_sncc_reduce(_lib.ptr(dev_items), n_items, _lib.ptr(partials), _lib.ptr(red), st),
               "fvr_sncc_reduce")


def target_energy(t: torch.Tensor) -> float:
    """sum |t|^2 in float64 of a (small) target spectrum, on the host."""
    x = t.detach().to("cpu").numpy().astype(np.complex128)
    return float(np.sum(x.real * x.real + x.imag * x.imag))


def pair_moments(a: torch.Tensor, b: torch.Tensor) -> np.ndarray:
    """[Re sum a b*, |a|^2, |b|^2, sum uu Im(a b*), sum vv Im(a b*)] of two (h, w) complex slices."""
    _lib.require_device(a, "pair_moments")
    _lib.require_device(b, "pair_moments")
    if a.shape != b.shape or a.dim() != 2:
        raise ValueError(f"pair_moments: shapes {tuple(a.shape)} and {tuple(b.shape)} must be equal 2D")
    a = a.to(torch.complex64).contiguous()
    b = b.to(torch.complex64).contiguous()
    out = torch.empty(5, dtype=torch.float64, device=a.device)
    _lib.check(_lib.lib().fvr_c64_pair_moments(a.data_ptr(), b.data_ptr(), a.shape[0], a.shape[1], _lib.ptr(out),
                                               _lib.stream_ptr(a.device)), "fvr_c64_pair_moments")
    return out.cpu().numpy()


def ncc_from_moments(A: float, E: float, T2: float) -> float:
    """NCC = A / (sqrt(E) sqrt(T2)), 0 if a norm is 0 (diff_fvr.py:190-195)."""
    ns, nt = math.sqrt(max(E, 0.0)), math.sqrt(max(T2, 0.0))
    if ns * nt == 0:
        return 0.0
    return A / (ns * nt)


class BiplanarEvaluator:
    """Batched convention-N biplanar loss + exact 6-DoF gradient.

    Loss per pose: L = 1/2 (1 - NCC_AP) + 1/2 (1 - NCC_LAT) with AP in-plane
    translation (tx, tz) and LAT (ty, tz) (gradients.py:76-108).

    ``evaluate`` runs three stream-ordered launches with no host sync and no
    allocation once the workspace is sized (``prepare``): float64 pose set-up
    (``fvr_pn_build_items``), the fused kernel, and the per-pose fixed-order
    reduction + combine (``fvr_pn_reduce_combine``).  Tensor poses give device
    results (views of the workspace, overwritten by the next call); NumPy poses
    give NumPy results.  ``evaluate_host`` builds the items on the host and
    combines there (the cross-check of the device set-up).  ``plan`` (a
    ``LaunchPlan``) sets the tile height, tile order and persistence.
    """

    def __init__(self, packed: PackedSpectrum, target_ap: torch.Tensor, target_lat: torch.Tensor,
                 adjoint: bool = True, herm: Optional[bool] = None, plan: Optional[LaunchPlan] = None):
        self.packed = packed
        self.device = packed.device
        self.hw = {"ap": _items.slice_dims_n(packed.shape, "ap"), "lateral": _items.slice_dims_n(packed.shape, "lateral")}
        fold_ok = packed.has_herm() and all(_items.herm_ok(*self.hw[v]) for v in ("ap", "lateral"))
        # Hermitian fold: half the samples, same sums (the default whenever the layouts and shapes allow it)
        self.herm = fold_ok if herm is None else bool(herm)
        if self.herm and not fold_ok:
            raise ValueError("Hermitian fold needs the wrapped layouts and even slice sides with h <= w")
        self.targets = tuple(torch.zeros(self.hw[v], dtype=torch.complex64, device=self.device)
                             for v in ("ap", "lateral"))
        self.folded = tuple(torch.zeros(self.hw[v] + (4,), dtype=torch.float32, device=self.device)
                            for v in ("ap", "lateral")) if self.herm else None
        self.t2 = np.zeros(2)
        self.set_targets(target_ap, target_lat)
        self.flags = (_lib.ITEM_ADJOINT if adjoint else 0) | (_lib.ITEM_HERM if self.herm else 0)
        self.lat_perm = packed.layout_for("lateral", self.herm)[1]
        self.lat_swap = int(tuple(self.lat_perm) == (0, 2, 1))
        self.ws = _Workspace(self.device)
        self.last_moments: Optional[np.ndarray] = None
        self.set_plan(plan or LaunchPlan())

    def set_plan(self, plan: LaunchPlan) -> None:
        """Choose tile height / order / persistence (rebuilds the item templates; a no-op
        for the plan already set)."""
        if getattr(self, "plan", None) == plan:
            return
        if plan.order not in ("xcd", "target") or plan.persist < 0:
            raise ValueError(f"bad launch plan {plan}")
        if plan.exact_box and not self.flags & _lib.ITEM_HERM:
            raise ValueError("exact_box applies to the Hermitian-fold kernel only")
        if not _lib.lib().fvr_sncc_has_depth(self.flags, plan.tile_h, plan.depth):
            raise ValueError(f"no fused-kernel variant for flags {self.flags} with tile height {plan.tile_h}, "
                             f"depth {plan.depth}")
        self.plan = plan
        self.rows = plan.tile_h
        tgt = self.folded if self.herm else self.targets
        self.templates = _lib.as_struct(np.concatenate([
            _items.n_template(self.packed, "ap", tgt[0].data_ptr(), self.flags, self.rows),
            _items.n_template(self.packed, "lateral", tgt[1].data_ptr(), self.flags, self.rows)]),
            _lib.ITEM_DTYPE)
        # every address-forming field the device builder copies comes from these templates
        _items.validate_items(self.build(np.zeros((1, 6)))[0], len(self.packed.descs), self.packed.descs, self.rows)
        self.view_tiles = (_items.tiles_for(*self.hw["ap"], self.rows, self.herm),
                           _items.tiles_for(*self.hw["lateral"], self.rows, self.herm))
        ta, tl = self.view_tiles
        self.tiles_per_pose = ta[0] * ta[1] + tl[0] * tl[1]
        self._maps: Dict[int, Tuple[Optional[torch.Tensor], int, int]] = {}
        self.plan_gen = getattr(self, "plan_gen", 0) + 1

    def state_key(self) -> Tuple:
        """Changes whenever something a captured HIP graph holds by address or by value
        changes: the plan (templates, tile maps), the targets' energies, the workspace."""
        return (self.plan_gen, self.target_gen, self.ws.version)

    def set_targets(self, target_ap: torch.Tensor, target_lat: torch.Tensor) -> None:
        """Copy new AP / LAT target spectra into the evaluator's buffers (a captured graph
        keeps the target energies it was captured with: capture again after this)."""
        for k, (name, t) in enumerate((("ap", target_ap), ("lateral", target_lat))):
            _lib.require_device(t, "BiplanarEvaluator target")
            if tuple(t.shape) != self.hw[name]:
                raise ValueError(f"{name} target shape {tuple(t.shape)} != slice shape {self.hw[name]}")
            self.targets[k].copy_(t)
            self.t2[k] = target_energy(self.targets[k])
            if self.folded is not None:
                self.folded[k].copy_(fold_target(self.targets[k]))
        self.target_gen = getattr(self, "target_gen", 0) + 1

    # -- device path ---------------------------------------------------------
    def prepare(self, n: int):
        """Size the workspace for n poses (no allocation in later calls of size <= n)."""
        ws = self.ws
        return (ws.get("poses", 6 * n, torch.float64),
                ws.get("items", 2 * n * _lib.ITEM_DTYPE.itemsize, torch.uint8),
                ws.get("jac", 2 * n * 18, torch.float64),
                ws.get("partials", n * self.tiles_per_pose * _lib.NMOM, torch.float32),
                ws.get("moments", 2 * n * _lib.NMOM, torch.float64),
                ws.get("loss", n, torch.float64),
                ws.get("grad", 6 * n, torch.float64))

    def launch_map(self, n: int):
        """(device tile map or None, n_slots, grid) of the plan for n poses (cached per n)."""
        got = self._maps.get(n)
        if got is None:
            n_tiles = n * self.tiles_per_pose
            if self.plan.order == "target":
                m = slot_map(target_major_entries(n, self.view_tiles))
                dev = torch.from_numpy(m.reshape(-1).copy()).to(self.device)
                n_slots = len(m)
            else:
                dev, n_slots = None, 8 * ((n_tiles + 7) // 8)
            grid = min(n_slots, 8 * ((self.plan.persist * N_CU + 7) // 8)) if self.plan.persist else 0
            got = (dev, n_slots, grid)
            self._maps[n] = got
        return got

    def build_device(self, poses: torch.Tensor, tile0: int = 0, items=None, jac=None) -> int:
        """Device pose set-up only: items + Jacobians into the workspace; returns N."""
        _lib.require_device(poses, "BiplanarEvaluator.build_device")
        if poses.dtype != torch.float64 or poses.dim() != 2 or poses.shape[1] != 6 or not poses.is_contiguous():
            raise ValueError("poses must be a contiguous (N, 6) float64 tensor")
        n = int(poses.shape[0])
        if n < 1:
            raise ValueError("empty pose batch")
        if items is None:
            _, items, jac = self.prepare(n)[:3]
            items, jac = _lib.ptr(items), _lib.ptr(jac)
        pz, py, px = self.packed.shape
        _lib.check(_lib.lib().fvr_pn_build_items(_lib.ptr(poses), n, self.templates.ctypes.data, float(pz),
                                                 float(py), float(px), self.lat_swap, int(tile0), items, jac,
                                                 _lib.stream_ptr(self.device)),
                   "fvr_pn_build_items")
        return n

    def evaluate_device(self, poses: torch.Tensor, kernel_events=None, want_moments: bool = True):
        """poses: contiguous (N, 6) float64 device tensor -> (loss (N,), grad (N, 6)) device views."""
        n = self.build_device(poses)
        _, items, jac, partials, moments, loss, grad = self.prepare(n)
        L = _lib.lib()
        st = _lib.stream_ptr(self.device)
        tmap, n_slots, grid = self.launch_map(n)
        if kernel_events is not None:
            kernel_events[0].record()
        _lib.check(L.fvr_sncc_eval_plan(_lib.ptr(self.packed.table), len(self.packed.descs), _lib.ptr(items), 2 * n,
                                        n * self.tiles_per_pose, _lib.ptr(partials),
                                        self.plan.variant(self.packed.code, self.flags),
                                        None if tmap is None else _lib.ptr(tmap), n_slots, grid, st), "fvr_sncc_eval")
        if kernel_events is not None:
            kernel_events[1].record()
        _lib.check(L.fvr_pn_reduce_combine(_lib.ptr(items), n, _lib.ptr(partials), _lib.ptr(jac), float(self.t2[0]),
                                           float(self.t2[1]), _lib.ptr(moments) if want_moments else None,
                                           _lib.ptr(loss), _lib.ptr(grad), st), "fvr_pn_reduce_combine")
        return loss[:n], grad[: 6 * n].view(n, 6)

    def evaluate_call(self, poses: torch.Tensor):
        """The same evaluation through the one-call C entry point ``fvr_pn_evaluate`` (set-up,
        fused kernel with the plan's tile height, depth and order -- "target" through the
        kernel's map-free target-major order of item pairs, persistence ignored -- reduce +
        combine; items, Jacobians and partials in a workspace sized by
        ``fvr_pn_evaluate_workspace``): (loss (N,), grad (N, 6)) device views."""
        _lib.require_device(poses, "BiplanarEvaluator.evaluate_call")
        if poses.dtype != torch.float64 or poses.dim() != 2 or poses.shape[1] != 6 or not poses.is_contiguous():
            raise ValueError("poses must be a contiguous (N, 6) float64 tensor")
        n = int(poses.shape[0])
        L = _lib.lib()
        nb = int(L.fvr_pn_evaluate_workspace(self.templates.ctypes.data, n))
        if nb < 0:
            raise ValueError("fvr_pn_evaluate_workspace rejected the templates")
        ws = self.ws.get("call_ws", nb, torch.uint8)
        loss = self.ws.get("call_loss", n, torch.float64)
        grad = self.ws.get("call_grad", 6 * n, torch.float64)
        pz, py, px = self.packed.shape
        plan = LaunchPlan(self.plan.tile_h, "xcd", 0, self.plan.depth, self.plan.exact_box)
        variant = plan.variant(self.packed.code, self.flags)
        if self.plan.order == "target":
            variant |= VARIANT_BIPLANAR_ORDER
        _lib.check(L.fvr_pn_evaluate(_lib.ptr(self.packed.table), len(self.packed.descs), self.templates.ctypes.data,
                                     float(pz), float(py), float(px), self.lat_swap,
                                     variant, _lib.ptr(poses), n, float(self.t2[0]),
                                     float(self.t2[1]), _lib.ptr(ws), nb, None, _lib.ptr(loss), _lib.ptr(grad),
                                     _lib.stream_ptr(self.device)), "fvr_pn_evaluate")
        return loss[:n], grad[: 6 * n].view(n, 6)

    def autotune_call(self, poses: torch.Tensor, reps: int = 5, use_cache: bool = True):
        """Autotune ``fvr_pn_evaluate`` on ``poses`` in the library (``fvr_autotune_pn_evaluate``)
        and set the matching plan: returns (plan, mean ms of the best variant).  The
        workspace is sized for the smallest tile height, so every candidate fits."""
        _lib.require_device(poses, "BiplanarEvaluator.autotune_call")
        if poses.dtype != torch.float64 or poses.dim() != 2 or poses.shape[1] != 6 or not poses.is_contiguous():
            raise ValueError("poses must be a contiguous (N, 6) float64 tensor")
        n = int(poses.shape[0])
        L = _lib.lib()
        t32 = self.templates.copy()
        t32["tile_h"] = 32
        nb = int(L.fvr_pn_evaluate_workspace(t32.ctypes.data, n))
        if nb < 0:
            raise ValueError("fvr_pn_evaluate_workspace rejected the templates")
        ws = self.ws.get("call_ws", nb, torch.uint8)
        loss = self.ws.get("call_loss", n, torch.float64)
        grad = self.ws.get("call_grad", 6 * n, torch.float64)
        pz, py, px = self.packed.shape
        best = ctypes.c_uint32(0)
        ms = ctypes.c_float(0.0)
        base = LaunchPlan(self.plan.tile_h, "xcd", 0, self.plan.depth, self.plan.exact_box)
        _lib.check(L.fvr_autotune_pn_evaluate(_lib.ptr(self.packed.table), len(self.packed.descs),
                                              self.templates.ctypes.data, float(pz), float(py), float(px),
                                              self.lat_swap, base.variant(self.packed.code, self.flags),
                                              _lib.ptr(poses), n, float(self.t2[0]), float(self.t2[1]), _lib.ptr(ws),
                                              nb, _lib.ptr(loss), _lib.ptr(grad), int(reps), int(bool(use_cache)),
                                              ctypes.byref(best), ctypes.byref(ms), _lib.stream_ptr(self.device)),
                   "fvr_autotune_pn_evaluate")
        v = int(best.value)
        plan = LaunchPlan((v >> 8) & 0xFF, "target" if v & VARIANT_BIPLANAR_ORDER else "xcd", 0, (v >> 16) & 0xF,
                          bool((v >> 20) & 1))
        self.set_plan(plan)
        return plan, float(ms.value)

    # -- two-launch loop form: begin() once, then step() per iteration ----------
    def begin(self, poses: torch.Tensor) -> int:
        """Build the items of ``poses`` (the first launch of a loop of ``step`` calls)."""
        n = self.build_device(poses)
        self.prepare(n)
        self.launch_map(n)
        return n

    def step(self, poses: torch.Tensor, adam=None, kernel_events=None):
        """One loop iteration in two launches: the fused kernel over the items built by the
        previous ``begin``/``step``, then ``fvr_pn_step`` (reduce + combine -> loss, grad at
        those poses; optionally one Adam step of ``poses`` in place (``adam``: a DeviceAdam);
        the items of the (updated) poses for the next call).  Returns (loss, grad) views."""
        n = int(poses.shape[0])
        _, items, jac, partials, moments, loss, grad = self.prepare(n)
        L = _lib.lib()
        st = _lib.stream_ptr(self.device)
        tmap, n_slots, grid = self.launch_map(n)
        if kernel_events is not None:
            kernel_events[0].record()
        _lib.check(L.fvr_sncc_eval_plan(_lib.ptr(self.packed.table), len(self.packed.descs), _lib.ptr(items), 2 * n,
                                        n * self.tiles_per_pose, _lib.ptr(partials),
                                        self.plan.variant(self.packed.code, self.flags),
                                        None if tmap is None else _lib.ptr(tmap), n_slots, grid, st), "fvr_sncc_eval")
        if kernel_events is not None:
            kernel_events[1].record()
        pz, py, px = self.packed.shape
        _pn_step(items, n, partials, jac, self.t2, None, loss, grad, poses, adam, self.templates, pz, py, px,
                 self.lat_swap, 0, st)
        return loss[:n], grad[: 6 * n].view(n, 6)

    def device_items(self, n: int) -> np.ndarray:
        """The items of the last device build (for cross-checks)."""
        raw = self.ws.bufs["items"][: 2 * n * _lib.ITEM_DTYPE.itemsize].cpu().numpy()
        return raw.view(_lib.ITEM_DTYPE).copy()

    def device_jacobians(self, n: int) -> np.ndarray:
        return self.ws.bufs["jac"][: 2 * n * 18].cpu().numpy().reshape(2 * n, 2, 3, 3).copy()

    def evaluate(self, poses, kernel_events=None):
        """poses (N, 6) (deg, voxels) -> (loss (N,), grad (N, 6)) float64.

        Device tensors in -> device tensors out (workspace views); NumPy in -> NumPy out.
        """
        if isinstance(poses, torch.Tensor) and poses.is_cuda:
            return self.evaluate_device(poses.to(torch.float64).contiguous(), kernel_events)
        P = np.ascontiguousarray(np.asarray(poses, dtype=np.float64).reshape(-1, 6))
        dev_poses = self.prepare(len(P))[0][: P.size].view(len(P), 6)
        dev_poses.copy_(torch.from_numpy(P))
        loss, grad = self.evaluate_device(dev_poses, kernel_events)
        self.last_moments = self.ws.bufs["moments"][: 2 * len(P) * _lib.NMOM].view(2 * len(P), _lib.NMOM).cpu().numpy()
        return loss.cpu().numpy(), grad.cpu().numpy()

    # -- host-built path (cross-check) -----------------------------------------
    def build(self, poses: np.ndarray):
        """Validated items and Jacobians for a (N, 6) pose batch (host, float64)."""
        items, jac, n_tiles = _items.n_biplanar_items(self.templates, self.packed.shape, poses, self.lat_perm,
                                                      self.rows)
        n_chk = _items.validate_items(items, len(self.packed.descs), self.packed.descs, self.rows)
        assert n_chk == n_tiles
        return items, jac, n_tiles

    def evaluate_host(self, poses):
        """Host-built items, fused kernel + per-item reduction, host float64 combine (NumPy out)."""
        if isinstance(poses, torch.Tensor):
            poses = poses.detach().to("cpu").numpy()
        poses = np.asarray(poses, dtype=np.float64).reshape(-1, 6)
        items, jac, n_tiles = self.build(poses)
        red = torch.empty((len(items), _lib.NMOM), dtype=torch.float64, device=self.device)
        hws = _Workspace(self.device)
        launch_items(self.packed, items, n_tiles, self.flags, hws, red)
        mom = red.cpu().numpy()
        return combine_biplanar(mom, jac, self.t2, self.hw)


def _pn_step(items, n, partials, jac, t2, moments, loss, grad, poses, adam, templates, pz, py, px, lat_swap, tile0, st,
             item0: int = 0, pose0: int = 0):
    """fvr_pn_step for poses [pose0, pose0 + n) whose items start at item 2 * pose0."""
    isz = _lib.ITEM_DTYPE.itemsize
    if adam is None:
        a = [None] * 7
    else:
        a = [adam.cfg.ctypes.data, _off(adam.m, 6 * pose0), _off(adam.v, 6 * pose0), _off(adam.t, pose0),
             _off(adam.done, pose0), _off(adam.best_loss, pose0), _off(adam.best_pose, 6 * pose0)]
    _lib.check(_lib.lib().fvr_pn_step(_off(items, 2 * pose0 * isz), n, _lib.ptr(partials), _off(jac, 2 * pose0 * 18),
                                      float(t2[0]), float(t2[1]), moments, _off(loss, pose0), _off(grad, 6 * pose0),
                                      _off(poses, 6 * pose0), *a, templates.ctypes.data, float(pz), float(py),
                                      float(px), lat_swap, tile0, st), "fvr_pn_step")


def _off(t: torch.Tensor, elems: int) -> ctypes.c_void_p:
    """Device address of element ``elems`` of a contiguous tensor."""
    return ctypes.c_void_p(t.data_ptr() + elems * t.element_size())


class RaggedBiplanarEvaluator:
    """Several volumes (each with its own targets and pose count) in ONE fused launch.

    ``segments``: BiplanarEvaluator per volume (same device, storage type and
    kernel flags); ``counts``: poses per volume.  Poses are one (sum counts, 6)
    float64 device tensor, volume k owning rows [offsets[k], offsets[k] + counts[k]).
    A step is: one device pose set-up per volume (its templates re-indexed into a
    combined layout table and its tiles counted from the running total), one
    fused launch over every item of every volume (ragged item sizes: the kernel
    maps each tile to its item), and one reduce/combine per volume.  No host
    sync and no allocation after ``prepare`` (graph-capturable).
    """

    def __init__(self, segments: Sequence[BiplanarEvaluator], counts: Sequence[int],
                 plan: Optional[LaunchPlan] = None):
        if len(segments) == 0 or len(segments) != len(counts) or any(int(c) < 1 for c in counts):
            raise ValueError("one positive pose count per segment")
        self.segments = list(segments)
        self.counts = [int(c) for c in counts]
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)[:-1]]).astype(int).tolist()
        self.n = int(sum(self.counts))
        s0 = self.segments[0]
        self.device = s0.device
        for s in self.segments:
            if s.device != self.device or s.packed.code != s0.packed.code or s.flags != s0.flags:
                raise ValueError("segments must share the device, the storage type and the kernel flags")
        self.flags = s0.flags
        self.code = s0.packed.code
        self.descs = np.concatenate([s.packed.descs for s in self.segments]).astype(_lib.LAYOUT_DTYPE)
        self.table = _lib.struct_to_device(self.descs, self.device, _lib.LAYOUT_DTYPE)
        self.lay0 = np.concatenate([[0], np.cumsum([len(s.packed.descs) for s in self.segments])[:-1]]).astype(int)
        self.ws = _Workspace(self.device)
        self.set_plan(plan or s0.plan)

    def set_plan(self, plan: LaunchPlan) -> None:
        """Set the plan of every segment and rebuild the combined tile map (a no-op for the
        plan already set)."""
        if getattr(self, "plan", None) == plan:
            return
        for s in self.segments:
            s.set_plan(plan)
        self.plan = plan
        self.plan_gen = getattr(self, "plan_gen", 0) + 1
        self.rows = plan.tile_h
        self.templates = []
        self.tile0 = []
        t = 0
        for k, s in enumerate(self.segments):
            tm = s.templates.copy()
            tm["layout"] += self.lay0[k]
            # the re-indexed templates against the combined table (the device builder copies them)
            probe, _, _ = _items.n_biplanar_items(tm, s.packed.shape, np.zeros((1, 6)), s.lat_perm, self.rows, t)
            _items.validate_items(probe, len(self.descs), self.descs, self.rows, tile0=t)
            self.templates.append(_lib.as_struct(tm, _lib.ITEM_DTYPE))
            self.tile0.append(t)
            t += self.counts[k] * s.tiles_per_pose
        self.n_tiles = t
        if t >= 2 ** 31:
            raise ValueError("ragged batch exceeds the 32-bit tile index")
        if plan.order == "target":
            ent = np.concatenate([target_major_entries(c, s.view_tiles, 2 * o)
                                  for s, c, o in zip(self.segments, self.counts, self.offsets)])
            m = slot_map(ent)
            self.map = torch.from_numpy(m.reshape(-1).copy()).to(self.device)
            self.n_slots = len(m)
        else:
            self.map, self.n_slots = None, 8 * ((t + 7) // 8)
        self.grid = min(self.n_slots, 8 * ((plan.persist * N_CU + 7) // 8)) if plan.persist else 0
        self.prepare()

    def prepare(self):
        ws, n = self.ws, self.n
        return (ws.get("items", 2 * n * _lib.ITEM_DTYPE.itemsize, torch.uint8),
                ws.get("jac", 2 * n * 18, torch.float64),
                ws.get("partials", self.n_tiles * _lib.NMOM, torch.float32),
                ws.get("loss", n, torch.float64),
                ws.get("grad", 6 * n, torch.float64))

    def evaluate_device(self, poses: torch.Tensor, kernel_events=None):
        """poses (sum counts, 6) contiguous float64 device tensor -> (loss (N,), grad (N, 6)) device views."""
        _lib.require_device(poses, "RaggedBiplanarEvaluator")
        if poses.dtype != torch.float64 or tuple(poses.shape) != (self.n, 6) or not poses.is_contiguous():
            raise ValueError(f"poses must be a contiguous ({self.n}, 6) float64 tensor")
        items, jac, partials, loss, grad = self.prepare()
        L = _lib.lib()
        st = _lib.stream_ptr(self.device)
        isz = _lib.ITEM_DTYPE.itemsize
        for k, s in enumerate(self.segments):
            o, c = self.offsets[k], self.counts[k]
            pz, py, px = s.packed.shape
            _lib.check(L.fvr_pn_build_items(_off(poses, 6 * o), c, self.templates[k].ctypes.data, float(pz),
                                            float(py), float(px), s.lat_swap, self.tile0[k], _off(items, 2 * o * isz),
                                            _off(jac, 2 * o * 18), st), "fvr_pn_build_items")
        if kernel_events is not None:
            kernel_events[0].record()
        _lib.check(L.fvr_sncc_eval_plan(_lib.ptr(self.table), len(self.descs), _lib.ptr(items), 2 * self.n,
                                        self.n_tiles, _lib.ptr(partials), self.plan.variant(self.code, self.flags),
                                        None if self.map is None else _lib.ptr(self.map), self.n_slots, self.grid,
                                        st), "fvr_sncc_eval")
        if kernel_events is not None:
            kernel_events[1].record()
        for k, s in enumerate(self.segments):
            o, c = self.offsets[k], self.counts[k]
            _lib.check(L.fvr_pn_reduce_combine(_off(items, 2 * o * isz), c, _lib.ptr(partials), _off(jac, 2 * o * 18),
                                               float(s.t2[0]), float(s.t2[1]), None, _off(loss, o),
                                               _off(grad, 6 * o), st), "fvr_pn_reduce_combine")
        return loss[: self.n], grad[: 6 * self.n].view(self.n, 6)

    def begin(self, poses: torch.Tensor) -> int:
        """Build every segment's items (the first launch of a loop of ``step`` calls)."""
        self._check_poses(poses)
        items, jac = self.prepare()[:2]
        L = _lib.lib()
        st = _lib.stream_ptr(self.device)
        isz = _lib.ITEM_DTYPE.itemsize
        for k, s in enumerate(self.segments):
            o, c = self.offsets[k], self.counts[k]
            pz, py, px = s.packed.shape
            _lib.check(L.fvr_pn_build_items(_off(poses, 6 * o), c, self.templates[k].ctypes.data, float(pz),
                                            float(py), float(px), s.lat_swap, self.tile0[k], _off(items, 2 * o * isz),
                                            _off(jac, 2 * o * 18), st), "fvr_pn_build_items")
        return self.n

    def step(self, poses: torch.Tensor, adam=None):
        """One loop iteration: one fused launch over every volume's items, then one
        ``fvr_pn_step`` per volume (reduce + combine, optional Adam, next items)."""
        self._check_poses(poses)
        items, jac, partials, loss, grad = self.prepare()
        L = _lib.lib()
        st = _lib.stream_ptr(self.device)
        _lib.check(L.fvr_sncc_eval_plan(_lib.ptr(self.table), len(self.descs), _lib.ptr(items), 2 * self.n,
                                        self.n_tiles, _lib.ptr(partials), self.plan.variant(self.code, self.flags),
                                        None if self.map is None else _lib.ptr(self.map), self.n_slots, self.grid,
                                        st), "fvr_sncc_eval")
        for k, s in enumerate(self.segments):
            pz, py, px = s.packed.shape
            _pn_step(items, self.counts[k], partials, jac, s.t2, None, loss, grad, poses, adam, self.templates[k], pz,
                     py, px, s.lat_swap, self.tile0[k], st, pose0=self.offsets[k])
        return loss[: self.n], grad[: 6 * self.n].view(self.n, 6)

    def state_key(self) -> Tuple:
        """Changes whenever something a captured HIP graph holds changes (the plan, the tile
        map, any segment's targets, the combined workspace)."""
        return (self.plan_gen, self.ws.version, tuple(s.target_gen for s in self.segments))

    def _check_poses(self, poses: torch.Tensor) -> None:
        _lib.require_device(poses, "RaggedBiplanarEvaluator")
        if poses.dtype != torch.float64 or tuple(poses.shape) != (self.n, 6) or not poses.is_contiguous():
            raise ValueError(f"poses must be a contiguous ({self.n}, 6) float64 tensor")

    def device_items(self) -> np.ndarray:
        raw = self.ws.bufs["items"][: 2 * self.n * _lib.ITEM_DTYPE.itemsize].cpu().numpy()
        return raw.view(_lib.ITEM_DTYPE).copy()


def combine_biplanar(mom: np.ndarray, jac: np.ndarray, t2, hw):
    """L = 1/2(1 - NCC_AP) + 1/2(1 - NCC_LAT) and its exact 6-DoF gradient (gradients.py:95-108).

    mom: (2N, 16) moments (AP, LAT per pose); jac: (2N, 2, 3, 3) du/dtheta,
    dv/dtheta per degree in layout axes; t2: target energies; hw: slice shapes.
    dNCC = Re<dS,T>/(|S||T|) - NCC Re<dS,S>/|S|^2 with dS/da = -2 pi i uu/W S.
    """
    N = len(mom) // 2
    loss = np.zeros(N)
    grad = np.zeros((N, 6))
    for view, name in ((0, "ap"), (1, "lateral")):
        m = mom[view::2]
        J = jac[view::2]
        H, W = hw[name]
        nt = math.sqrt(t2[view])
        ns = np.sqrt(np.maximum(m[:, 1], 0.0))
        den = ns * nt
        ok = den != 0.0
        inv = np.where(ok, 1.0 / np.where(ok, den, 1.0), 0.0)
        ncc = m[:, 0] * inv
        G = np.einsum("na,naj->nj", m[:, 4:7], J[:, 0]) + np.einsum("na,naj->nj", m[:, 7:10], J[:, 1])
        Hm = np.einsum("na,naj->nj", m[:, 10:13], J[:, 0]) + np.einsum("na,naj->nj", m[:, 13:16], J[:, 1])
        E = np.where(ok, m[:, 1], 1.0)
        d_rot = np.where(ok[:, None], G * inv[:, None] - ncc[:, None] * Hm / E[:, None], 0.0)
        d_a = TWO_PI / W * m[:, 2] * inv
        d_b = TWO_PI / H * m[:, 3] * inv
        loss += 0.5 * (1.0 - ncc)
        grad[:, :3] -= 0.5 * d_rot
        grad[:, 3 if view == 0 else 4] -= 0.5 * d_a
        grad[:, 5] -= 0.5 * d_b
    return loss, grad
