"""Algorithmic byte count of the fused slice kernel (for roofline reporting).

For one convention-N slice, U is the number of distinct spectrum entries with
a non-zero trilinear weight (scipy box rule); for scaled storage the number
of distinct 8^3 scale bricks is counted too.  Bytes per slice are
``U * elem_bytes + bricks * 4``; the target spectrum (complex64, shared by all
candidate poses of a batch) is added once per launch by the caller.
Runs on whatever device the tensors live on (torch ops; a measurement helper,
not part of the rendering path).
"""

from __future__ import annotations

from typing import Tuple

import numpy as np
import torch

from ._items import euler_to_basis_n, slice_dims_n


def slice_footprint(pad_shape, pose, view: str, device="cpu") -> Tuple[int, int]:
    """(U, bricks) of one convention-N slice at pose (yaw, pitch, roll in degrees)."""
    pz, py, px = pad_shape
    h, w = slice_dims_n(pad_shape, view)
    u, v = euler_to_basis_n(pose[0], pose[1], pose[2], view)
    uu = torch.arange(w, dtype=torch.float64, device=device) - w / 2.0
    vv = torch.arange(h, dtype=torch.float64, device=device) - h / 2.0
    V, U_ = torch.meshgrid(vv, uu, indexing="ij")
    c = np.array(pad_shape, dtype=np.float64) / 2.0
    p = torch.stack([U_ * u[a] + V * v[a] + c[a] for a in range(3)], dim=-1).reshape(-1, 3)
    n = torch.tensor([pz, py, px], dtype=torch.float64, device=device)
    inb = ((p >= 0) & (p <= n - 1)).all(dim=1)
    p = p[inb]
    i0 = torch.floor(p).to(torch.int64)
    t = p - i0
    keys = []
    bkeys = []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wgt = ((t[:, 0] if dz else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1])
                       * (t[:, 2] if dx else 1 - t[:, 2]))
                idx = i0 + torch.tensor([dz, dy, dx], device=device)
                idx = idx[wgt > 0]
                keys.append((idx[:, 0] * py + idx[:, 1]) * px + idx[:, 2])
                # storage bricks: storage index = logical + 1, 8^3 bricks
                b = (idx + 1) // 8
                bkeys.append((b[:, 0] * 4096 + b[:, 1]) * 4096 + b[:, 2])
    U = int(torch.unique(torch.cat(keys)).numel())
    B = int(torch.unique(torch.cat(bkeys)).numel())
    return U, B


def batch_algorithmic_bytes(pad_shape, poses: np.ndarray, elem_bytes: int, scaled: bool, device="cpu") -> dict:
    """Algorithmic bytes of one biplanar launch over ``poses`` (AP + LAT per pose)."""
    total_u = 0
    total_b = 0
    for p in poses:
        for view in ("ap", "lateral"):
            U, B = slice_footprint(pad_shape, p, view, device)
            total_u += U
            total_b += B
    h_ap, w_ap = slice_dims_n(pad_shape, "ap")
    h_la, w_la = slice_dims_n(pad_shape, "lateral")
    targ = (h_ap * w_ap + h_la * w_la) * 8
    spec = total_u * elem_bytes + (total_b * 4 if scaled else 0)
    return {"unique_entries": total_u, "bricks": total_b, "spectrum_bytes": spec, "target_bytes": targ,
            "bytes": spec + targ, "samples": len(poses) * (h_ap * w_ap + h_la * w_la)}
