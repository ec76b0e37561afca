"""Pose-recovery error metric (drop-in for ``src/fvr/registration_3d2d.py:112-175``)."""

from __future__ import annotations

from typing import Tuple

import numpy as np

from ._items import euler_rotation_n


def compute_mtre_6dof(true_pose: Tuple[float, ...], est_pose: Tuple[float, ...], volume_shape: Tuple[int, int, int],
                      spacing: Tuple[float, float, float] = (1.0, 1.0, 1.0), translation_units: str = "voxels") -> float:
    """Mean target registration error in mm over eight synthetic landmarks at r = min(shape)/4."""
    spacing = np.asarray(spacing, dtype=float)
    if spacing.shape != (3,) or not np.all(np.isfinite(spacing)) or np.any(spacing <= 0):
        raise ValueError("spacing must contain three finite positive values in mm.")
    if translation_units not in ("voxels", "mm"):
        raise ValueError("translation_units must be 'voxels' or 'mm'.")
    nz, ny, nx = volume_shape
    c = np.array([nz / 2.0, ny / 2.0, nx / 2.0])
    r = min(nz, ny, nx) * 0.25
    corners = np.array([[c[0] + sz * r, c[1] + sy * r, c[2] + sx * r]
                        for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)])

    def transform(pts, pose):
        yaw, pitch, roll, tx, ty, tz = pose
        rot = euler_rotation_n(yaw, pitch, roll)
        tr = np.array([tz, ty, tx], dtype=float)
        if translation_units == "voxels":
            tr = tr * spacing
        return ((pts - c) * spacing) @ rot.T + tr

    return float(np.mean(np.linalg.norm(transform(corners, true_pose) - transform(corners, est_pose), axis=1)))
