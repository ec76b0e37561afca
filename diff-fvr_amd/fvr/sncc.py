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
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from . import _items, _lib
from .spectrum import PackedSpectrum

TWO_PI = 2.0 * math.pi


def tile_rows() -> int:
    return int(_lib.lib().fvr_sncc_tile_rows())


class _Workspace:
    """Grow-only device buffers keyed by name (no allocation once sized)."""

    def __init__(self, device):
        self.device = device
        self.bufs: Dict[str, torch.Tensor] = {}

    def get(self, name: str, numel: int, dtype) -> torch.Tensor:
        t = self.bufs.get(name)
        if t is None or t.numel() < numel or t.dtype != dtype:
            t = torch.empty(max(numel, 1), dtype=dtype, device=self.device)
            self.bufs[name] = t
        return t


def item_tiles(items: np.ndarray, rows: int) -> np.ndarray:
    tx = (items["w"] + 63) // 64
    ty = (items["h"] + rows - 1) // rows
    return (tx * ty).astype(np.int64)


def run_items(packed: PackedSpectrum, items: np.ndarray, flags: int,
              ws: Optional[_Workspace] = None) -> torch.Tensor:
    """Evaluate host-built items; returns (n_items, 16) float64 moments on the device.

    ``items['tile_base']`` is (re)assigned here.  Every item must carry the
    same ``flags`` (one kernel variant per launch).
    """
    L = _lib.lib()
    rows = tile_rows()
    items = np.array(items, copy=True)
    ntile = item_tiles(items, rows)
    items["tile_base"] = np.concatenate([[0], np.cumsum(ntile)[:-1]]).astype(np.int32)
    n_tiles = int(ntile.sum())
    n_items = len(items)
    ws = ws or _Workspace(packed.device)
    dev_items = _lib.struct_to_device(items, packed.device)
    partials = ws.get("partials", n_tiles * _lib.NMOM, torch.float32)
    red = torch.empty((n_items, _lib.NMOM), dtype=torch.float64, device=packed.device)
    st = _lib.stream_ptr(packed.device)
    variant = packed.code | (flags << 4)
    _lib.check(L.fvr_sncc_eval_v(_lib.ptr(packed.table), _lib.ptr(dev_items), n_items, n_tiles,
                                 _lib.ptr(partials), variant, st), "fvr_sncc_eval")
    _lib.check(L.fvr_sncc_reduce(_lib.ptr(dev_items), n_items, _lib.ptr(partials), _lib.ptr(red), st),
               "fvr_sncc_reduce")
    return red


def c64_norm2(t: torch.Tensor) -> torch.Tensor:
    """sum |t|^2 (float64 device scalar, fixed order) of a complex64 device tensor."""
    _lib.require_device(t, "c64_norm2")
    t = t.to(torch.complex64).contiguous()
    out = torch.empty(1, dtype=torch.float64, device=t.device)
    _lib.check(_lib.lib().fvr_c64_norm2(t.data_ptr(), t.numel(), _lib.ptr(out), _lib.stream_ptr(t.device)),
               "fvr_c64_norm2")
    return out


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
    """Batched convention-N biplanar loss + exact 6-DoF gradient on the device.

    Loss per pose: L = 1/2 (1 - NCC_AP) + 1/2 (1 - NCC_LAT) with AP in-plane
    translation (tx, tz) and LAT (ty, tz) (gradients.py:76-108); gradient by
    the fused adjoint (rotations exact, translations in closed form).
    """

    def __init__(self, packed: PackedSpectrum, target_ap: torch.Tensor, target_lat: torch.Tensor,
                 adjoint: bool = True):
        self.packed = packed
        self.device = packed.device
        pz, py, px = packed.shape
        self.pad = torch.tensor([pz, py, px], dtype=torch.float64, device=self.device)
        self.hw = {"ap": _items.slice_dims_n(packed.shape, "ap"), "lateral": _items.slice_dims_n(packed.shape, "lateral")}
        self.targets = tuple(torch.zeros(self.hw[v], dtype=torch.complex64, device=self.device)
                             for v in ("ap", "lateral"))
        self.t2 = torch.zeros(2, dtype=torch.float64, device=self.device)
        self.set_targets(target_ap, target_lat)
        self.slice_hw = torch.tensor([*self.hw["ap"], *self.hw["lateral"]], dtype=torch.float64, device=self.device)
        self.rows = tile_rows()
        self.flags = _lib.ITEM_ADJOINT if adjoint else 0
        tm = np.concatenate([
            _items.n_template(packed, "ap", self.targets[0].data_ptr(), self.flags, self.rows),
            _items.n_template(packed, "lateral", self.targets[1].data_ptr(), self.flags, self.rows)])
        self.templates = _lib.struct_to_device(tm, self.device)
        self.lat_transposed = int("zxy" in packed.index)
        ta = _items.tiles_for(*self.hw["ap"], self.rows)
        tl = _items.tiles_for(*self.hw["lateral"], self.rows)
        self.tiles_per_pose = ta[0] * ta[1] + tl[0] * tl[1]
        self.ws = _Workspace(self.device)

    def set_targets(self, target_ap: torch.Tensor, target_lat: torch.Tensor) -> None:
        """Copy new AP / LAT target spectra into the evaluator's buffers (stream-ordered)."""
        for k, (name, t) in enumerate((("ap", target_ap), ("lateral", target_lat))):
            _lib.require_device(t, "BiplanarEvaluator target")
            if tuple(t.shape) != self.hw[name]:
                raise ValueError(f"{name} target shape {tuple(t.shape)} != slice shape {self.hw[name]}")
            self.targets[k].copy_(t)
            _lib.check(_lib.lib().fvr_c64_norm2(self.targets[k].data_ptr(), self.targets[k].numel(),
                                                ctypes.c_void_p(self.t2.data_ptr() + 8 * k),
                                                _lib.stream_ptr(self.device)), "fvr_c64_norm2")

    def n_tiles(self, n_poses: int) -> int:
        return n_poses * self.tiles_per_pose

    def prepare(self, n_poses: int) -> None:
        """Size the workspace for n_poses (call before graph capture)."""
        self._bufs(n_poses)

    def _bufs(self, n: int):
        ws = self.ws
        items = ws.get("items", 2 * n * _lib.ITEM_DTYPE.itemsize, torch.uint8)
        jac = ws.get("jac", 2 * n * 18, torch.float64)
        partials = ws.get("partials", self.n_tiles(n) * _lib.NMOM, torch.float32)
        red = ws.get("red", 2 * n * _lib.NMOM, torch.float64)
        loss = ws.get("loss", n, torch.float64)
        grad = ws.get("grad", 6 * n, torch.float64)
        return items, jac, partials, red, loss, grad

    def evaluate(self, poses: torch.Tensor, out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
                 kernel_events=None):
        """poses (N, 6) float64 device tensor (deg, voxels) -> (loss (N,), grad (N, 6)) float64.

        ``kernel_events`` = (start, end) torch.cuda.Events recorded on the
        launch stream around the fused kernel alone (for timing it).
        """
        _lib.require_device(poses, "BiplanarEvaluator.evaluate")
        if poses.dtype != torch.float64 or poses.dim() != 2 or poses.shape[1] != 6 or not poses.is_contiguous():
            raise ValueError("poses must be a contiguous (N, 6) float64 tensor")
        n = int(poses.shape[0])
        L = _lib.lib()
        items, jac, partials, red, loss, grad = self._bufs(n)
        if out is not None:
            loss, grad = out
        st = _lib.stream_ptr(self.device)
        _lib.check(L.fvr_pn_build_items(_lib.ptr(poses), n, _lib.ptr(self.templates), _lib.ptr(self.pad),
                                        _lib.ptr(items), _lib.ptr(jac), self.lat_transposed, st), "fvr_pn_build_items")
        if kernel_events is not None:
            kernel_events[0].record()
        _lib.check(L.fvr_sncc_eval_v(_lib.ptr(self.packed.table), _lib.ptr(items), 2 * n, self.n_tiles(n),
                                     _lib.ptr(partials), self.packed.code | (self.flags << 4), st), "fvr_sncc_eval")
        if kernel_events is not None:
            kernel_events[1].record()
        _lib.check(L.fvr_sncc_reduce(_lib.ptr(items), 2 * n, _lib.ptr(partials), _lib.ptr(red), st), "fvr_sncc_reduce")
        _lib.check(L.fvr_pn_combine(_lib.ptr(red), _lib.ptr(jac), n, _lib.ptr(self.t2), _lib.ptr(self.slice_hw),
                                    _lib.ptr(loss), _lib.ptr(grad), st), "fvr_pn_combine")
        return loss[:n], grad[:6 * n].view(n, 6)

    def moments(self, n: int) -> torch.Tensor:
        """Per-item moments of the last evaluate() (2n x 16, AP then LAT per pose)."""
        return self.ws.bufs["red"][: 2 * n * _lib.NMOM].view(2 * n, _lib.NMOM)
