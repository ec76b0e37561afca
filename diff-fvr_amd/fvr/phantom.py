"""Seeded synthetic CT phantoms (inputs for benchmarks and tests).

``ellipsoid_phantom``: a sum of randomly oriented ellipsoids with attenuation
U(0.01, 0.05)/mm on a zero background, centres in the central 60 % of the
cube and semi-axes U(0.05 n, 0.25 n).  Pure NumPy, generated slab by slab so
that 256^3 fits in modest host memory; identical on every machine for a seed.
"""

from __future__ import annotations

import numpy as np


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def ellipsoid_phantom(n: int = 96, n_ellipsoids: int = 12, seed: int = 1234,
                      mu_range=(0.01, 0.05), slab: int = 16) -> np.ndarray:
    """(n, n, n) float32 volume (Z, Y, X) in 1/mm."""
    rng = np.random.default_rng(seed)
    params = []
    for _ in range(n_ellipsoids):
        centre = rng.uniform(-0.3 * n, 0.3 * n, 3)
        axes = rng.uniform(0.05 * n, 0.25 * n, 3)
        rot = _random_rotation(rng)
        mu = rng.uniform(*mu_range)
        params.append((centre, axes, rot, mu))
    vol = np.zeros((n, n, n), dtype=np.float64)
    ax = np.arange(n, dtype=np.float64) - (n - 1) / 2.0

    def span(c, r):  # index range of the axis coordinates within [c - r, c + r]
        lo = int(np.searchsorted(ax, c - r, side="left"))
        hi = int(np.searchsorted(ax, c + r, side="right"))
        return lo, hi

    for centre, axes, rot, mu in params:
        r = float(axes.max())  # bounding sphere of the ellipsoid
        (y0, y1), (x0, x1) = span(centre[1], r), span(centre[2], r)
        zlo, zhi = span(centre[0], r)
        for z0 in range(zlo, zhi, slab):
            z1 = min(zhi, z0 + slab)
            zz, yy, xx = np.meshgrid(ax[z0:z1], ax[y0:y1], ax[x0:x1], indexing="ij")
            d = (np.stack([zz, yy, xx], axis=-1) - centre) @ rot  # coordinates in the ellipsoid frame
            vol[z0:z1, y0:y1, x0:x1] += mu * (((d / axes) ** 2).sum(-1) <= 1.0)
    return vol.astype(np.float32)


def random_poses(n: int, rot_deg: float, trans: float, seed: int) -> np.ndarray:
    """(n, 6) float64 poses: rotations U(-rot_deg, rot_deg), translations U(-trans, trans)."""
    rng = np.random.default_rng(seed)
    p = np.empty((n, 6))
    p[:, :3] = rng.uniform(-rot_deg, rot_deg, (n, 3))
    p[:, 3:] = rng.uniform(-trans, trans, (n, 3))
    return p
