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
    rows = tile_rows()
    items = _lib.as_struct(items, _lib.ITEM_DTYPE).copy()
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
    variant = packed.code | (flags << 4)
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
    translation (tx, tz) and LAT (ty, tz) (gradients.py:76-108).  Per call:
    float64 pose set-up on the host (vectorised), one fused HIP launch over all
    2N items, the fixed-order float64 reduction on the device, and the float64
    biplanar combine on the host from the 16 moments per item.
    """

    def __init__(self, packed: PackedSpectrum, target_ap: torch.Tensor, target_lat: torch.Tensor,
                 adjoint: bool = True):
        self.packed = packed
        self.device = packed.device
        self.hw = {"ap": _items.slice_dims_n(packed.shape, "ap"), "lateral": _items.slice_dims_n(packed.shape, "lateral")}
        self.targets = tuple(torch.zeros(self.hw[v], dtype=torch.complex64, device=self.device)
                             for v in ("ap", "lateral"))
        self.t2 = np.zeros(2)
        self.set_targets(target_ap, target_lat)
        self.rows = tile_rows()
        self.flags = _lib.ITEM_ADJOINT if adjoint else 0
        self.templates = _lib.as_struct(np.concatenate([
            _items.n_template(packed, "ap", self.targets[0].data_ptr(), self.flags, self.rows),
            _items.n_template(packed, "lateral", self.targets[1].data_ptr(), self.flags, self.rows)]),
            _lib.ITEM_DTYPE)
        self.lat_perm = packed.layout_for("lateral")[1]
        self.ws = _Workspace(self.device)
        self.last_moments: Optional[np.ndarray] = None

    def set_targets(self, target_ap: torch.Tensor, target_lat: torch.Tensor) -> None:
        """Copy new AP / LAT target spectra into the evaluator's buffers."""
        for k, (name, t) in enumerate((("ap", target_ap), ("lateral", target_lat))):
            _lib.require_device(t, "BiplanarEvaluator target")
            if tuple(t.shape) != self.hw[name]:
                raise ValueError(f"{name} target shape {tuple(t.shape)} != slice shape {self.hw[name]}")
            self.targets[k].copy_(t)
            self.t2[k] = target_energy(self.targets[k])

    def build(self, poses: np.ndarray):
        """Validated items and Jacobians for a (N, 6) pose batch (host, float64)."""
        items, jac, n_tiles = _items.n_biplanar_items(self.templates, self.packed.shape, poses, self.lat_perm,
                                                      self.rows)
        n_chk = _items.validate_items(items, len(self.packed.descs), self.packed.descs, self.rows)
        assert n_chk == n_tiles
        return items, jac, n_tiles

    def launch(self, items: np.ndarray, n_tiles: int, kernel_events=None) -> torch.Tensor:
        """Fused kernel + reduction for prebuilt items; returns the device moments (2N, 16)."""
        red = self.ws.get("red", len(items) * _lib.NMOM, torch.float64)[: len(items) * _lib.NMOM]
        red = red.view(len(items), _lib.NMOM)
        if kernel_events is not None:
            # time the fused kernel alone: wrap eval between events, reduction after
            L = _lib.lib()
            dev_items = self.ws.get("items", len(items) * _lib.ITEM_DTYPE.itemsize, torch.uint8)
            dev_items[: len(items) * _lib.ITEM_DTYPE.itemsize].copy_(
                torch.from_numpy(_lib.struct_bytes(items, _lib.ITEM_DTYPE)))
            partials = self.ws.get("partials", n_tiles * _lib.NMOM, torch.float32)
            st = _lib.stream_ptr(self.device)
            kernel_events[0].record()
            _lib.check(L.fvr_sncc_eval_v(_lib.ptr(self.packed.table), len(self.packed.descs), _lib.ptr(dev_items),
                                         len(items), n_tiles, _lib.ptr(partials),
                                         self.packed.code | (self.flags << 4), st), "fvr_sncc_eval")
            kernel_events[1].record()
            _lib.check(L.fvr_sncc_reduce(_lib.ptr(dev_items), len(items), _lib.ptr(partials), _lib.ptr(red), st),
                       "fvr_sncc_reduce")
        else:
            launch_items(self.packed, items, n_tiles, self.flags, self.ws, red)
        return red

    def combine(self, mom: np.ndarray, jac: np.ndarray):
        """float64 loss (N,) and gradient (N, 6) from item moments (host)."""
        return combine_biplanar(mom, jac, self.t2, self.hw)

    def evaluate(self, poses, kernel_events=None):
        """poses (N, 6) (deg, voxels; NumPy or tensor) -> (loss (N,), grad (N, 6)) float64 NumPy."""
        if isinstance(poses, torch.Tensor):
            poses = poses.detach().to("cpu").numpy()
        poses = np.asarray(poses, dtype=np.float64).reshape(-1, 6)
        items, jac, n_tiles = self.build(poses)
        red = self.launch(items, n_tiles, kernel_events)
        mom = red.cpu().numpy()
        self.last_moments = mom
        return self.combine(mom, jac)


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
