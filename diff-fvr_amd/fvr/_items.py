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
# Every sample coordinate is biased by +2^-22 voxel (and the convention-N box
# edge with it).  Stepping with rounded 32.32 steps can leave a coordinate
# that is exactly on a lattice plane in float64 (identity poses, grid end
# points) up to ~2^-23 voxel below it; the bias puts such samples on the plane,
# so floor(), the box test and the one-sided derivative there agree with the
# float64 / float32 computations of the reference.
SNAP_FX = 1 << 10


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


def tiles_for(h: int, w: int, tile_rows: int, herm: bool = False) -> Tuple[int, int]:
    """(tiles per row, tile rows) of an item: tiles of 64 columns x tile_rows rows; a
    Hermitian-fold item covers h/2 rows in tiles of 32 columns x 2 tile_rows rows."""
    if herm:
        return (w + 31) // 32, (h // 2 + 2 * tile_rows - 1) // (2 * tile_rows)
    return (w + 63) // 64, (h + tile_rows - 1) // tile_rows


def guard_offset(desc) -> int:
    """Byte offset of the all-zero guard plane (storage plane sdim0 - 2) of a layout."""
    return (int(desc["sdim"][0]) - 2) * int(desc["plane_bytes"])


def herm_ok(h: int, w: int) -> bool:
    """Slice shapes the Hermitian-fold kernel covers: even sides, h <= w (column 0 of the
    lower rows is evaluated by the idle lanes of row h/2)."""
    return h % 2 == 0 and w % 2 == 0 and h <= w and h >= 2


def n_item(packed, view: str, yaw, pitch, roll, a, b, target_ptr: int, flags: int, tile_base: int,
           tile_rows: int, slice_out: int = 0) -> np.ndarray:
    """One convention-N item (numpy structured scalar array of length 1).  With
    ``ITEM_HERM`` in ``flags`` the item uses the wrapped layout and ``target_ptr`` is
    the folded (T, conj T(mirror)) float4 array of the view."""
    pad_shape = packed.shape
    herm = bool(flags & _lib.ITEM_HERM)
    lay_idx, perm = packed.layout_for(view, herm)
    desc = packed.desc(lay_idx)
    h, w = slice_dims_n(pad_shape, view)
    u, v = euler_to_basis_n(yaw, pitch, roll, view)
    uu0, vv0 = -w / 2.0, -h / 2.0
    c = np.array(pad_shape, dtype=np.float64) / 2.0
    cg = c + uu0 * u + vv0 * v
    it = np.zeros(1, dtype=_lib.ITEM_DTYPE)
    it["c"] = to_fx(cg[list(perm)]) + SNAP_FX
    it["ustep"] = to_fx(u[list(perm)])
    it["vstep"] = to_fx(v[list(perm)])
    n = np.array(desc["n"], dtype=np.int64)
    it["box_hi"] = ((n - 1) << 32).astype(np.uint64) + np.uint64(SNAP_FX)
    it["ph_col"] = turns_fx(a / w)
    it["ph_col0"] = turns_fx(uu0 * a / w)
    it["ph_row"] = turns_fx(b / h)
    it["ph_row0"] = turns_fx(vv0 * b / h)
    it["cw0"], it["cws"], it["rw0"], it["rws"] = uu0, 1.0, vv0, 1.0
    it["h"], it["w"] = h, w
    it["layout"] = lay_idx
    it["lo_int"] = 1
    it["guard_off"] = guard_offset(desc)
    it["tile_base"] = tile_base
    it["tiles_x"] = tiles_for(h, w, tile_rows, herm)[0]
    it["tile_h"] = tile_rows
    it["flags"] = flags
    it["target"] = target_ptr
    it["slice_out"] = slice_out
    return it


def n_template(packed, view: str, target_ptr: int, flags: int, tile_rows: int) -> np.ndarray:
    """Static fields of a convention-N item (geometry filled per pose by n_biplanar_items)."""
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


# ---------------------------------------------------------------------------
# Vectorised convention-N set-up for batches of poses (AP + LAT per pose)
# ---------------------------------------------------------------------------
def _rot_and_derivs(poses: np.ndarray):
    """R (N,3,3) = Rz(yaw) Ry(roll) Rx(pitch) and dR/dtheta (N,3,3,3) per DEGREE."""
    y, p, r = (np.asarray(poses[:, k], dtype=np.float64) * DEG for k in range(3))
    N = len(y)
    z0, o1 = np.zeros(N), np.ones(N)

    def m(rows):
        return np.stack([np.stack(row, axis=-1) for row in rows], axis=-2)

    cy, sy, cp, sp, cr, sr = np.cos(y), np.sin(y), np.cos(p), np.sin(p), np.cos(r), np.sin(r)
    Rz = m([[o1, z0, z0], [z0, cy, -sy], [z0, sy, cy]])
    Ry = m([[cr, z0, sr], [z0, o1, z0], [-sr, z0, cr]])
    Rx = m([[cp, -sp, z0], [sp, cp, z0], [z0, z0, o1]])
    dRz = m([[z0, z0, z0], [z0, -sy, -cy], [z0, cy, -sy]])
    dRy = m([[-sr, z0, cr], [z0, z0, z0], [-cr, z0, -sr]])
    dRx = m([[-sp, -cp, z0], [cp, -sp, z0], [z0, z0, z0]])
    R = Rz @ Ry @ Rx
    dR = np.stack([dRz @ Ry @ Rx, Rz @ Ry @ dRx, Rz @ dRy @ Rx], axis=1) * DEG  # (N, 3, 3, 3)
    return R, dR


def n_biplanar_items(templates: np.ndarray, pad_shape, poses: np.ndarray, lat_perm, tile_rows: int,
                     tile0: int = 0):
    """Items (2N, AP then LAT per pose) and Jacobians (2N, 2, 3, 3) for convention N
    (tile_base counted from ``tile0``).

    ``templates`` are the AP / LAT items whose static fields are kept.  The
    geometry restates diff_fvr.py:56-150 with tx on AP, ty on LAT and tz as the
    vertical in-plane translation of both views (gradients.py:89-95).
    """
    poses = np.asarray(poses, dtype=np.float64).reshape(-1, 6)
    N = len(poses)
    R, dR = _rot_and_derivs(poses)
    pz, py, px = pad_shape
    c = np.array([pz / 2.0, py / 2.0, px / 2.0])
    templates = _lib.as_struct(templates, _lib.ITEM_DTYPE)
    items = np.empty(2 * N, dtype=_lib.ITEM_DTYPE)
    jac = np.empty((2 * N, 2, 3, 3))
    tiles = [int(np.prod(tiles_for(int(t["h"]), int(t["w"]), tile_rows, bool(t["flags"] & _lib.ITEM_HERM))))
             for t in templates]
    for view in (0, 1):
        t = templates[view]
        perm = [0, 1, 2] if view == 0 else list(lat_perm)
        ub = 2 if view == 0 else 1
        H, W = float(pz), float(px if view == 0 else py)
        uu0, vv0 = -W / 2.0, -H / 2.0
        u = R[:, :, ub]  # rot @ u_base, (N, 3) in (z, y, x)
        v = R[:, :, 0]
        cg = c[None, :] + uu0 * u + vv0 * v
        sl = items[view::2]
        sl[:] = t
        sl["c"] = to_fx(cg[:, perm]) + SNAP_FX
        sl["ustep"] = to_fx(u[:, perm])
        sl["vstep"] = to_fx(v[:, perm])
        a = poses[:, 3] if view == 0 else poses[:, 4]
        b = poses[:, 5]
        sl["ph_col"] = turns_fx(a / W)
        sl["ph_col0"] = turns_fx(uu0 * a / W)
        sl["ph_row"] = turns_fx(b / H)
        sl["ph_row0"] = turns_fx(vv0 * b / H)
        sl["tile_base"] = tile0 + np.arange(N) * (tiles[0] + tiles[1]) + (tiles[0] if view == 1 else 0)
        # du/dtheta_j = dR_j @ u_base, dv/dtheta_j = dR_j @ v_base, layout axes
        jac[view::2, 0] = np.transpose(dR[:, :, :, ub], (0, 2, 1))[:, perm, :]
        jac[view::2, 1] = np.transpose(dR[:, :, :, 0], (0, 2, 1))[:, perm, :]
    return items, jac, N * (tiles[0] + tiles[1])


def validate_items(items: np.ndarray, n_layouts: int, descs: np.ndarray, tile_rows: int, tile0: int = 0) -> int:
    """Host-side checks of a launch's descriptors; returns the tile count
    (``tile0``: tile_base of the first item, for a segment of a larger launch).

    Every field the kernel uses to form an address is checked here, so a
    malformed descriptor raises before anything reaches the device.
    """
    if len(items) == 0:
        raise ValueError("empty item batch")
    if items.dtype != _lib.ITEM_DTYPE or items.dtype.itemsize != _lib.ITEM_DTYPE.itemsize:
        raise ValueError("items must use the C item layout (fvr._lib.ITEM_DTYPE)")
    lay = items["layout"]
    if np.any(lay < 0) or np.any(lay >= n_layouts):
        raise ValueError("item layout index out of range")
    if np.any(items["h"] <= 0) or np.any(items["w"] <= 0):
        raise ValueError("item slice shape must be positive")
    if tile_rows not in (32, 64, 128) or np.any(items["tile_h"] != tile_rows):
        raise ValueError(f"item tile_h must be the launch's tile height ({tile_rows}: 32, 64 or 128)")
    if np.any((items["lo_int"] != 0) & (items["lo_int"] != 1)):
        raise ValueError("item lo_int must be 0 or 1")
    if np.any(items["target"] == 0):
        raise ValueError("item target pointer is null")
    if np.any(((items["flags"] & _lib.ITEM_WRITE_SLICE) != 0) & (items["slice_out"] == 0)):
        raise ValueError("slice output requested without a buffer")
    herm = (items["flags"] & _lib.ITEM_HERM) != 0
    rows = np.where(herm, items["h"] // 2, items["h"])
    tw, th = np.where(herm, 32, 64), np.where(herm, 2 * tile_rows, tile_rows)
    ntile = ((items["w"] + tw - 1) // tw) * ((rows + th - 1) // th)
    if np.any(herm):
        hi_ = items[herm]
        if np.any((hi_["flags"] & (_lib.ITEM_HP | _lib.ITEM_WRITE_SLICE)) != 0) or np.any(hi_["lo_int"] != 1):
            raise ValueError("Hermitian-fold items are convention N without high-pass or slice output")
        if np.any(hi_["h"] % 2) or np.any(hi_["w"] % 2) or np.any(hi_["h"] > hi_["w"]):
            raise ValueError("Hermitian-fold items need even h <= w")
        if np.any((descs[hi_["layout"]]["flags"] & _lib.LAYOUT_WRAP) == 0):
            raise ValueError("Hermitian-fold items need a wrapped layout")
        # the fold samples up to floor(p) = n: its corners (+lo, +1) must stay inside the storage
        dh = descs[hi_["layout"]]
        if np.any((hi_["box_hi"] >> np.uint64(32)).astype(np.int64) + 3 > dh["sdim"].astype(np.int64) - 1):
            raise ValueError("Hermitian-fold item box exceeds the wrapped layout's storage")
    if np.any(items["tiles_x"] != (items["w"] + tw - 1) // tw):
        raise ValueError("item tiles_x inconsistent with its width")
    base = int(tile0) + np.concatenate([[0], np.cumsum(ntile)[:-1]])
    if not np.array_equal(items["tile_base"].astype(np.int64), base):
        raise ValueError("item tile_base not the running tile count")
    d = descs[lay]
    lo = items["lo_int"].astype(np.int64)
    hi = (items["box_hi"] >> np.uint64(32)).astype(np.int64)  # (n_items, 3)
    # inside samples have floor(p) <= hi; their corners (+lo, +1) must stay inside the storage
    if np.any(hi + lo[:, None] + 1 > d["sdim"].astype(np.int64) - 1):
        raise ValueError("item box exceeds the layout's storage")
    s_g = d["sdim"][:, 0].astype(np.int64) - 2
    if np.any(items["guard_off"].astype(np.int64) != s_g * d["plane_bytes"]):
        raise ValueError("item guard offset is not the layout's zero plane")
    return int(ntile.sum())


# ---------------------------------------------------------------------------
# Convention T (torch_engine.py:138-234): standard xyz Euler angles, a
# half-bin grid, per-corner zero padding, optional high-pass weight
# ---------------------------------------------------------------------------
T_VIEW_LAYOUT = {"ap": "zxy", "lateral": "zyx"}  # lanes (slice columns) walk the fast axis


def t_rotation_and_derivs(poses: np.ndarray):
    """R (N,3,3) = Rz(yaw) Ry(pitch) Rx(roll) (standard xyz, torch_engine.py:138-174) and
    dR/dtheta (N,3,3,3) per DEGREE for theta = (yaw, pitch, roll)."""
    yz, py_, rx = (np.asarray(poses[:, k], dtype=np.float64) * DEG for k in range(3))
    N = len(yz)
    z0, o1 = np.zeros(N), np.ones(N)

    def m(rows):
        return np.stack([np.stack(row, axis=-1) for row in rows], axis=-2)

    cz, sz, cy, sy, cx, sx = np.cos(yz), np.sin(yz), np.cos(py_), np.sin(py_), np.cos(rx), np.sin(rx)
    Rz = m([[cz, -sz, z0], [sz, cz, z0], [z0, z0, o1]])
    Ry = m([[cy, z0, sy], [z0, o1, z0], [-sy, z0, cy]])
    Rx = m([[o1, z0, z0], [z0, cx, -sx], [z0, sx, cx]])
    dRz = m([[-sz, -cz, z0], [cz, -sz, z0], [z0, z0, z0]])
    dRy = m([[-sy, z0, cy], [z0, z0, z0], [-cy, z0, -sy]])
    dRx = m([[z0, z0, z0], [z0, -sx, -cx], [z0, cx, -sx]])
    R = Rz @ Ry @ Rx
    dR = np.stack([dRz @ Ry @ Rx, Rz @ dRy @ Rx, Rz @ Ry @ dRx], axis=1) * DEG
    return R, dR


_RZ90 = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])


def t_center_scale(pad_shape) -> np.ndarray:
    """(cx, cy, cz) = ((W-1)/2, (H-1)/2, (D-1)/2) of the padded volume (torch_engine.py:200-205)."""
    D, H, W = pad_shape
    return np.array([0.5 * max(W - 1, 1), 0.5 * max(H - 1, 1), 0.5 * max(D - 1, 1)], dtype=np.float64)


def t_view_axes(poses: np.ndarray, view: str, pad_shape):
    """u_xyz, v_xyz (N,3) and their per-degree derivatives (N,3 angles,3) for a view."""
    R, dR = t_rotation_and_derivs(np.asarray(poses, dtype=np.float64).reshape(-1, 6))
    if view == "lateral":
        R = R @ _RZ90
        dR = dR @ _RZ90
    s = t_center_scale(pad_shape)
    return R[:, 0, :] * s, R[:, 1, :] * s, dR[:, :, 0, :] * s, dR[:, :, 1, :] * s


def t_grid(out_res: int) -> Tuple[float, float]:
    """First value and step of linspace(-1, 1, out_res) (torch_engine.py:115-118)."""
    return -1.0, (2.0 / (out_res - 1) if out_res > 1 else 0.0)


def t_phase_span(out_res: int, spacing0: float) -> float:
    """The phase-ramp denominator out_res * spacing + 1e-6 (torch_engine.py:227-234)."""
    return out_res * float(spacing0) + 1e-6


def t_template(packed, view: str, out_res: int, target_ptr: int, flags: int, tile_rows: int,
               hp_sigma: float | None = None) -> np.ndarray:
    """Static fields of a convention-T item: layout, box, guard plane, grid sizes, high-pass weights."""
    lay_name = T_VIEW_LAYOUT[view]
    lay_idx = packed.index[lay_name]
    desc = packed.desc(lay_idx)
    it = np.zeros(1, dtype=_lib.ITEM_DTYPE)
    n = np.array(desc["n"], dtype=np.int64)
    # per-corner zero padding: coordinates shifted by +1 (lo_int = 0), inside when -1 < p < n
    # (the edge is not moved with the bias: floor() must stay <= n)
    it["box_hi"] = (((n + 1) << 32) - 1).astype(np.uint64)
    f0, df = t_grid(out_res)
    it["cw0"], it["cws"], it["rw0"], it["rws"] = f0, df, f0, df
    if hp_sigma is not None:
        flags |= _lib.ITEM_HP
        it["hp_inv2s2"] = 1.0 / (2.0 * float(hp_sigma) ** 2)
        step = out_res / (out_res - 1) if out_res > 1 else 0.0
        it["hp_x0"], it["hp_dx"], it["hp_y0"], it["hp_dy"] = -out_res / 2.0, step, -out_res / 2.0, step
    it["h"], it["w"] = out_res, out_res
    it["layout"] = lay_idx
    it["lo_int"] = 0
    it["guard_off"] = guard_offset(desc)
    it["tiles_x"] = tiles_for(out_res, out_res, tile_rows)[0]
    it["tile_h"] = tile_rows
    it["flags"] = flags
    it["target"] = target_ptr
    return it


def t_items(template: np.ndarray, pad_shape, poses: np.ndarray, view: str, out_res: int, spacing0: float,
            tile_rows: int, target_ptrs=None):
    """Items (N,) and Jacobians (N, 2, 3, 3) of convention T for one view.

    Sample (row i, col j) sits at ``center + fx_i u + fy_j v`` with
    ``fx = fy = linspace(-1, 1, out_res)`` (rows follow u, columns follow v:
    torch_engine.py:115-118, 196-211); phase ``exp(-2 pi i (fx_i tx + fy_j ty) / span)``.
    ``jac[:, 0]`` = dv/dtheta (paired with the column moments), ``jac[:, 1]`` =
    du/dtheta (row moments), layout axes x angles, per degree.
    """
    poses = np.asarray(poses, dtype=np.float64).reshape(-1, 6)
    N = len(poses)
    t = _lib.as_struct(template, _lib.ITEM_DTYPE)[0]
    perm = [0, 2, 1] if T_VIEW_LAYOUT[view] == "zxy" else [0, 1, 2]
    u, v, du, dv = t_view_axes(poses, view, pad_shape)
    cen = t_center_scale(pad_shape)
    f0, df = t_grid(out_res)
    zyx = [2, 1, 0]  # xyz -> (z, y, x)
    order = [zyx[a] for a in perm]  # layout axis -> xyz component
    c0 = cen[None, :] + f0 * u + f0 * v + 1.0  # +1: per-corner coordinates (lo_int = 0)
    items = np.empty(N, dtype=_lib.ITEM_DTYPE)
    items[:] = t
    items["c"] = to_fx(c0[:, order]) + SNAP_FX
    items["ustep"] = to_fx(df * v[:, order])  # along a row: next column
    items["vstep"] = to_fx(df * u[:, order])  # down a column: next row
    span = t_phase_span(out_res, spacing0)
    tx, ty = poses[:, 3], poses[:, 4]
    items["ph_row0"] = turns_fx(f0 * tx / span)
    items["ph_row"] = turns_fx(df * tx / span)
    items["ph_col0"] = turns_fx(f0 * ty / span)
    items["ph_col"] = turns_fx(df * ty / span)
    ntile = tiles_for(out_res, out_res, tile_rows)
    items["tile_base"] = np.arange(N) * (ntile[0] * ntile[1])
    if target_ptrs is not None:
        items["target"] = np.asarray(target_ptrs, dtype=np.uint64)
    jac = np.empty((N, 2, 3, 3))
    jac[:, 0] = np.transpose(dv[:, :, order], (0, 2, 1))
    jac[:, 1] = np.transpose(du[:, :, order], (0, 2, 1))
    return items, jac, N * ntile[0] * ntile[1]
