"""Algorithmic byte count of the fused slice kernel (for roofline reporting).

For one convention-N slice, U is the number of distinct spectrum entries with
a non-zero trilinear weight (scipy box rule).  Bytes per slice are
``U * elem_bytes`` (plus the 4-byte scale of scaled storage once per launch); the
target spectrum (complex64, shared by all candidate poses of a batch) is added once
per launch.
Runs on whatever device the tensors live on (torch ops; a measurement helper,
not part of the rendering path).
"""

from __future__ import annotations

from typing import Tuple

import numpy as np
import torch

from ._items import euler_to_basis_n, slice_dims_n


def _herm_slots(h: int, w: int, device):
    """(ii, jj, mirror) of the Hermitian-fold schedule (sncc_herm_kernel): rows h/2+1..h-1,
    row h/2 right of the centre plus columns 0 and w/2, column 0 of rows 1..h/2-1, row 0."""
    hh, hw = h // 2, w // 2
    I, J = torch.meshgrid(torch.arange(hh + 1, h, device=device), torch.arange(w, device=device), indexing="ij")
    ii, jj, mm = [I.reshape(-1)], [J.reshape(-1)], [(J >= 1).reshape(-1)]
    jr = torch.arange(w, device=device)
    keep = (jr == 0) | (jr >= hw)
    ii.append(torch.full_like(jr[keep], hh)), jj.append(jr[keep]), mm.append(jr[keep] > hw)
    rm = torch.arange(1, min(hw, hh), device=device)
    ii.append(rm), jj.append(torch.zeros_like(rm)), mm.append(torch.zeros_like(rm, dtype=torch.bool))
    ii.append(torch.zeros_like(jr)), jj.append(jr), mm.append(torch.zeros_like(jr, dtype=torch.bool))
    return torch.cat(ii), torch.cat(jj), torch.cat(mm)


def slice_footprint(pad_shape, pose, view: str, device="cpu", herm: bool = False) -> Tuple[int, int]:
    """(U, samples) of one convention-N slice at pose (yaw, pitch, roll in degrees).

    ``herm``: the samples the Hermitian-fold kernel evaluates (about half of the slice,
    each inside if it or its mirror is inside the scipy box) on the wrapped lattice.
    """
    pz, py, px = pad_shape
    h, w = slice_dims_n(pad_shape, view)
    u, v = euler_to_basis_n(pose[0], pose[1], pose[2], view)
    c = np.array(pad_shape, dtype=np.float64) / 2.0
    n = torch.tensor([pz, py, px], dtype=torch.float64, device=device)
    if herm:
        ii, jj, mirror = _herm_slots(h, w, device)
        U_, V = jj.to(torch.float64) - w / 2.0, ii.to(torch.float64) - h / 2.0
        p = torch.stack([U_ * u[a] + V * v[a] + c[a] for a in range(3)], dim=-1)
        q = n - p
        inb = ((p >= 0) & (p <= n - 1)).all(dim=1) | (mirror & ((q >= 0) & (q <= n - 1)).all(dim=1))
    else:
        uu = torch.arange(w, dtype=torch.float64, device=device) - w / 2.0
        vv = torch.arange(h, dtype=torch.float64, device=device) - h / 2.0
        V, U_ = torch.meshgrid(vv, uu, indexing="ij")
        p = torch.stack([U_ * u[a] + V * v[a] + c[a] for a in range(3)], dim=-1).reshape(-1, 3)
        inb = ((p >= 0) & (p <= n - 1)).all(dim=1)
    samples = int(p.shape[0])
    p = p[inb]
    i0 = torch.floor(p).to(torch.int64)
    t = p - i0
    keys = []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wgt = ((t[:, 0] if dz else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1])
                       * (t[:, 2] if dx else 1 - t[:, 2]))
                idx = i0 + torch.tensor([dz, dy, dx], device=device)
                idx = idx[wgt > 0]
                keys.append((idx[:, 0] * (py + 2) + idx[:, 1]) * (px + 2) + idx[:, 2])
    U = int(torch.unique(torch.cat(keys)).numel())
    return U, samples


def batch_algorithmic_bytes(pad_shape, poses: np.ndarray, elem_bytes: int, scaled: bool, device="cpu",
                            herm: bool = False) -> dict:
    """Algorithmic bytes of one biplanar launch over ``poses`` (AP + LAT per pose).

    ``herm``: counted for the Hermitian-fold schedule (its half-plane entries and its
    16-byte folded target entries)."""
    total_u = total_s = 0
    for p in poses:
        for view in ("ap", "lateral"):
            U, S = slice_footprint(pad_shape, p, view, device, herm)
            total_u += U
            total_s += S
    h_ap, w_ap = slice_dims_n(pad_shape, "ap")
    h_la, w_la = slice_dims_n(pad_shape, "lateral")
    targ = (total_s // max(len(poses), 1)) * 16 if herm else (h_ap * w_ap + h_la * w_la) * 8
    spec = total_u * elem_bytes + (4 if scaled else 0)
    return {"unique_entries": total_u, "spectrum_bytes": spec, "target_bytes": targ,
            "bytes": spec + targ, "samples": total_s, "schedule": "hermitian-fold" if herm else "full-slice"}


def pou_slice_footprint(P: int, u_xyz, v_xyz, device="cpu") -> int:
    """Distinct block-spectrum entries one PoU slab-image slice reads (the same for every
    block of a pose): the P x P samples P/2 + (j - P/2) u + (i - P/2) v of the per-corner
    trilinear stencil (cone_beam_fvr.py:480-505; u, v in (x, y, z) storage order), corners
    inside [0, P - 1]^3 with a non-zero weight."""
    k = torch.arange(P, dtype=torch.float64, device=device) - P // 2
    I, J = torch.meshgrid(k, k, indexing="ij")
    u = torch.as_tensor(np.ascontiguousarray(u_xyz, np.float64), device=device)
    v = torch.as_tensor(np.ascontiguousarray(v_xyz, np.float64), device=device)
    p = (P // 2) + J.reshape(-1, 1) * u + I.reshape(-1, 1) * v  # (K, 3) in (x, y, z)
    i0 = torch.floor(p).to(torch.int64)
    t = p - i0
    keys = []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                d = torch.tensor([dx, dy, dz], device=device)
                wgt = ((t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1])
                       * (t[:, 2] if dz else 1 - t[:, 2]))
                idx = i0 + d
                ok = (wgt > 0) & ((idx >= 0) & (idx <= P - 1)).all(dim=1)
                idx = idx[ok]
                keys.append((idx[:, 2] * P + idx[:, 1]) * P + idx[:, 0])
    return int(torch.unique(torch.cat(keys)).numel())


def pou_texels_mask(model, pose_mats, lattice):
    """(B, N, P, P) bool: the slab-image texels that the pinhole warp reads with a non-zero
    bilinear weight, and the warp's arguments (d_b, cu, cv, rx, ry, ic).  The image gradient
    of ``pou_pinhole_warp_sum`` for an all-ones output gradient is positive exactly there
    (every bilinear weight and 1/cos is >= 0).  With a 256^2 detector at 0.4 mm per pixel at
    the isocentre only part of each block's P x P image falls inside the field of view."""
    from .triton_slice import pou_pinhole_warp_sum
    dev = model.F_blk_ri.device
    R, t = pose_mats[:, :3, :3].to(dev), pose_mats[:, :3, 3].to(dev)
    src0, _, u0, v0, n0 = (x.to(dev) for x in lattice[:5])
    srcs = torch.einsum("nij,j->ni", R, src0) + t
    u_det, v_det, n_0 = (torch.einsum("nij,j->ni", R, a) for a in (u0, v0, n0))
    d_b, cu, cv = (x.detach() for x in model._compute_block_offsets(srcs, u_det, v_det, n_0))
    rx, ry, ic = (x.to(dev) for x in lattice[5:8])
    N, P = int(R.shape[0]), int(model.P)
    imgs = torch.ones((model.B, N, P, P), device=dev, dtype=torch.float32, requires_grad=True)
    out = pou_pinhole_warp_sum(imgs, d_b, cu, cv, rx, ry, ic, inv_ds=1.0 / model.ds_block, c_idx=float(P // 2))
    out.backward(torch.ones_like(out))
    return imgs.grad > 0, (d_b, cu, cv, rx, ry, ic)


def pou_texels_read(model, pose_mats, lattice) -> int:
    """Number of texels in ``pou_texels_mask``, summed over (block, pose)."""
    return int(pou_texels_mask(model, pose_mats, lattice)[0].sum())


def pou_step_bytes(B: int, P: int, K: int, u_xyz: np.ndarray, v_xyz: np.ndarray, elem_bytes: int,
                   texels_read: int = None, device="cpu") -> dict:
    """Algorithmic bytes of one PoU forward + adjoint step over N poses (u, v: (N, 3) in
    (x, y, z)), each byte counted once per kernel that must move it:
      slice + inverse FFT   U elements per (block, pose) read (the inverse FFT needs the whole
                            slice), the slab-image texels the warp reads written;
      warp                  those texels read, the K-pixel float32 projection written;
      fused adjoint         U elements re-read, those texels read, the K-pixel output gradient read.
    ``texels_read`` (float32 texels over all (block, pose), from ``pou_texels_read``); None
    counts every texel of the P x P images.  The ray cache (3 K floats) is shared by every
    pose and counted once per kernel."""
    U = [pou_slice_footprint(P, u, v, device) for u, v in zip(u_xyz, v_xyz)]
    N = len(U)
    gather = sum(U) * B * elem_bytes
    img = (N * B * P * P if texels_read is None else int(texels_read)) * 4
    rays = 3 * K * 4
    fwd = gather + img
    warp = img + N * K * 4 + rays
    adj = gather + img + N * K * 4 + rays
    return {"unique_entries_per_pose_mean": float(np.mean(U)), "slab_texels_read": img // 4,
            "slice_ifft": fwd, "warp_fwd": warp, "render_bwd": adj, "step": fwd + warp + adj}
