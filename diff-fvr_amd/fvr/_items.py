"""Host-side work-item descriptors for the fused Spectral-NCC kernel.

A work item is one (pose, view) slice evaluation.  Geometry is set up in
float64 and handed to the kernel as 32.32 fixed-point sample coordinates, so
the per-sample arithmetic on the device is exact integer stepping (see the
header of ``csrc/fvr_sncc.hip``).

Convention N (``diff_fvr.py:106-150``): sample (row i, col j) of a slice of
shape (H, W) sits at ``c + uu u + vv v`` with ``uu = j - W/2``,
``vv = i - H/2``, ``c = (pz/2, py/2, px/2)``, and is interpolated with scipy's
order-1 'constant' rule (the box test ``0 <= p <= n-1``).
"""

from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from . import _lib

FX = float(2 ** 32)
DEG = math.pi / 180.0


def to_fx(x) -> np.ndarray:
    return np.rint(np.asarray(x, dtype=np.float64) * FX).astype(np.int64)


def turns_fx(x) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    x = x - np.floor(x)
    return (np.rint(x * FX).astype(np.int64) & 0xFFFFFFFF).astype(np.uint32)


# ---------------------------------------------------------------------------
# Convention N geometry (diff_fvr.py:56-104)
# ---------------------------------------------------------------------------
def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def _ry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def euler_rotation_n(yaw_deg, pitch_deg, roll_deg) -> np.ndarray:
    """rot = Rz(yaw) Ry(roll) Rx(pitch) in (Z, Y, X) index space (diff_fvr.py:64-87)."""
    return _rz(yaw_deg * DEG) @ _ry(roll_deg * DEG) @ _rx(pitch_deg * DEG)


VIEW_BASES = {
    "lateral": (np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0])),
    "ap": (np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])),
    "axial": (np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0])),
}


def view_key(view: str) -> str:
    return view if view in ("lateral", "ap") else "axial"


def euler_to_basis_n(yaw_deg, pitch_deg, roll_deg, view="lateral"):
    rot = euler_rotation_n(yaw_deg, pitch_deg, roll_deg)
    ub, vb = VIEW_BASES[view_key(view)]
    return rot @ ub, rot @ vb


def slice_dims_n(pad_shape, view) -> Tuple[int, int]:
    """(slice_h, slice_w) per view (diff_fvr.py:121-126)."""
    pz, py, px = pad_shape
    if view == "lateral":
        return pz, py
    if view == "ap":
        return pz, px
    return py, px


def tiles_for(h: int, w: int, tile_rows: int) -> Tuple[int, int]:
    return (w + 63) // 64, (h + tile_rows - 1) // tile_rows


def n_item(packed, view: str, yaw, pitch, roll, a, b, target_ptr: int, flags: int, tile_base: int,
           tile_rows: int, slice_out: int = 0) -> np.ndarray:
    """One convention-N item (numpy structured scalar array of length 1)."""
    pad_shape = packed.shape
    lay_idx, perm = packed.layout_for(view)
    desc = packed.desc(lay_idx)
    h, w = slice_dims_n(pad_shape, view)
    u, v = euler_to_basis_n(yaw, pitch, roll, view)
    uu0, vv0 = -w / 2.0, -h / 2.0
    c = np.array(pad_shape, dtype=np.float64) / 2.0
    cg = c + uu0 * u + vv0 * v
    it = np.zeros(1, dtype=_lib.ITEM_DTYPE)
    it["c"] = to_fx(cg[list(perm)])
    it["ustep"] = to_fx(u[list(perm)])
    it["vstep"] = to_fx(v[list(perm)])
    n = np.array(desc["n"], dtype=np.int64)
    it["box_hi"] = ((n - 1) << 32).astype(np.uint64)
    it["ph_col"] = turns_fx(a / w)
    it["ph_col0"] = turns_fx(uu0 * a / w)
    it["ph_row"] = turns_fx(b / h)
    it["ph_row0"] = turns_fx(vv0 * b / h)
    it["cw0"], it["cws"], it["rw0"], it["rws"] = uu0, 1.0, vv0, 1.0
    it["h"], it["w"] = h, w
    it["layout"] = lay_idx
    it["lo_int"] = 1
    it["guard_off"] = (int(desc["sdim"][0]) - 2) * int(desc["plane_bytes"])
    it["tile_base"] = tile_base
    it["tiles_x"] = tiles_for(h, w, tile_rows)[0]
    it["flags"] = flags
    it["target"] = target_ptr
    it["slice_out"] = slice_out
    return it


def n_template(packed, view: str, target_ptr: int, flags: int, tile_rows: int) -> np.ndarray:
    """Static fields of a convention-N item (geometry filled on the device)."""
    return n_item(packed, view, 0.0, 0.0, 0.0, 0.0, 0.0, target_ptr, flags, 0, tile_rows)


def n_jacobian(yaw, pitch, roll, view, perm) -> Tuple[np.ndarray, np.ndarray]:
    """du/dtheta, dv/dtheta (3x3, per degree) in layout axes for convention N."""
    y, p, r = yaw * DEG, pitch * DEG, roll * DEG
    c_, s_ = math.cos, math.sin

    def drz(a):
        return np.array([[0.0, 0.0, 0.0], [0.0, -s_(a), -c_(a)], [0.0, c_(a), -s_(a)]])

    def dry(a):
        return np.array([[-s_(a), 0.0, c_(a)], [0.0, 0.0, 0.0], [-c_(a), 0.0, -s_(a)]])

    def drx(a):
        return np.array([[-s_(a), -c_(a), 0.0], [c_(a), -s_(a), 0.0], [0.0, 0.0, 0.0]])

    Rz, Ry, Rx = _rz(y), _ry(r), _rx(p)
    ds = [drz(y) @ Ry @ Rx, Rz @ Ry @ drx(p), Rz @ dry(r) @ Rx]
    ub, vb = VIEW_BASES[view_key(view)]
    du = np.stack([d @ ub for d in ds], axis=1) * DEG
    dv = np.stack([d @ vb for d in ds], axis=1) * DEG
    return du[list(perm)], dv[list(perm)]
