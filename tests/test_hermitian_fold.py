"""The Hermitian fold of the batched convention-N kernel, checked on the CPU in float64.

The fused kernel's Hermitian mode evaluates half of each slice: a sample at p
also stands for its mirror n - p, whose value is the conjugate on the wrapped
lattice, with the scipy box rule applied to each of the two positions
separately, the unpaired row 0 / column 0 evaluated on their own.  This test
restates that sample schedule with the oracle's spectrum and interpolation and
checks that it reproduces all 16 moments of the full slice (the sums the
non-folded kernel and the oracle form), at poses whose slices cross the box
faces, so that the asymmetric edge of the box is exercised.
"""

import numpy as np
import pytest

from oracle import diff_fvr_np as O


def _trilinear_ext(F_ext, p):
    """Plain trilinear value and gradient on the wrapped lattice (zeros beyond it)."""
    p = np.asarray(p, np.float64)
    f = np.floor(p).astype(np.int64)
    t = p - f
    val = 0.0
    grad = np.zeros(3, dtype=complex)
    for c in range(8):
        d = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
        idx = tuple(f[a] + d[a] for a in range(3))
        inside = all(0 <= idx[a] < F_ext.shape[a] for a in range(3))
        v = F_ext[idx] if inside else 0.0
        w = [t[a] if d[a] else 1.0 - t[a] for a in range(3)]
        val += v * w[0] * w[1] * w[2]
        for a in range(3):
            s = 1.0 if d[a] else -1.0
            grad[a] += v * s * np.prod([w[b] for b in range(3) if b != a])
    return val, grad


def _fx(p):
    """Coordinates on the kernel's 32.32 fixed-point grid (float64 rounding noise of
    +-1e-16 at a box face would otherwise decide in / out differently for a sample and
    its mirror, as it can for the float64 reference itself)."""
    return np.rint(np.asarray(p) * 2.0 ** 32) / 2.0 ** 32


def _moments_full(F, pose, view, T, fold=False):
    """The 16 sums of the full slice (scipy box rule), by definition (``fold``: the
    oracle's Hermitian-fold derivative convention at lattice-exact coordinates)."""
    H, W = O.slice_dims(F.shape, view)
    u, v = O.euler_to_basis(*pose[:3], view=view)
    c = np.array(F.shape) / 2.0
    a = pose[3] if view == "ap" else pose[4]
    b = pose[5]
    ii, jj = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    uu, vv = (jj - W / 2).ravel(), (ii - H / 2).ravel()
    P = _fx(c[:, None] + uu[None, :] * u[:, None] + vv[None, :] * v[:, None])  # (3, K)
    if fold:
        vals, grads = O.trilinear_value_and_grad(F, P, left=O.fold_mirror_mask(H, W), wrap=True)
    else:
        vals, grads = O.trilinear_value_and_grad(F, P)
    m = np.zeros(16)
    for k in range(P.shape[1]):
        ph = np.exp(-2j * np.pi * (uu[k] * a / W + vv[k] * b / H))
        Wt = ph * np.conj(T.ravel()[k])
        _acc(m, complex(vals[k]), grads[:, k], Wt, 1.0, uu[k], vv[k])
    return m


def _acc(m, I, g, Wt, wE, cw, rw):
    IW = I * Wt
    m[0] += IW.real
    m[1] += wE * abs(I) ** 2
    m[2] += cw * IW.imag
    m[3] += rw * IW.imag
    for k in range(3):
        gk = (g[k] * Wt).real
        hk = wE * (g[k] * np.conj(I)).real
        m[4 + k] += cw * gk
        m[7 + k] += rw * gk
        m[10 + k] += cw * hk
        m[13 + k] += rw * hk


def _moments_fold(F, pose, view, T):
    """The kernel's Hermitian-fold schedule (sncc_herm_kernel) in float64."""
    n = np.array(F.shape)
    F_ext = np.zeros(tuple(n + 1), dtype=complex)
    F_ext[tuple(slice(0, k + 1) for k in n)] = F[np.ix_(*[np.arange(k + 1) % k for k in n])]
    H, W = O.slice_dims(F.shape, view)
    hh, hw = H // 2, W // 2
    u, v = O.euler_to_basis(*pose[:3], view=view)
    c = n / 2.0
    a = pose[3] if view == "ap" else pose[4]
    b = pose[5]

    def inbox(p):
        return bool(np.all(p >= 0) and np.all(p <= n - 1))

    slots = []
    for j in range(W):
        for i in range(hh + 1, H):
            slots.append((i, j, j >= 1))
        if 1 <= j < hw:
            if j <= hh - 1:
                slots.append((j, 0, False))  # column 0 of rows 1 .. h/2-1, on an idle lane of row h/2
        else:
            slots.append((hh, j, j > hw))
        slots.append((0, j, False))  # the unpaired row 0
    m = np.zeros(16)
    for ii, jj, mirror in slots:
        uu, vv = jj - W / 2, ii - H / 2
        p = _fx(c + uu * u + vv * v)
        bp = inbox(p)
        bq = mirror and inbox(n - p)
        if not (bp or bq):
            continue
        val, g = _trilinear_ext(F_ext, p)
        ph = np.exp(-2j * np.pi * (uu * a / W + vv * b / H))
        Tf = (T[ii, jj] if bp else 0.0) + (np.conj(T[H - ii, W - jj]) if bq else 0.0)
        _acc(m, val, g, ph * np.conj(Tf), float(bp) + float(bq), uu, vv)
    return m


@pytest.mark.parametrize("view", ["ap", "lateral"])
@pytest.mark.parametrize("pose", [(7.3, -4.1, 11.9, 1.3, -0.7, 0.4), (-14.6, 13.2, -9.7, -2.2, 1.9, -1.1)])
def test_fold_reproduces_full_slice_moments(view, pose):
    rng = np.random.default_rng(8)
    vol = np.zeros((6, 7, 8))
    vol[1:5, 2:6, 1:7] = rng.uniform(0.0, 1.0, (4, 4, 6))
    F = O.make_spectrum(vol)  # padded (12, 14, 16): AP 12 x 16, LAT 12 x 14 (even, h <= w)
    H, W = O.slice_dims(F.shape, view)
    T = rng.standard_normal((H, W)) + 1j * rng.standard_normal((H, W))
    pose = np.array(pose)
    full = _moments_full(F, pose, view, T)
    fold = _moments_fold(F, pose, view, T)
    scale = np.maximum(np.abs(full), 1e-3 * np.abs(full).max())
    np.testing.assert_allclose(fold, full, atol=1e-9 * np.abs(full).max(), rtol=1e-9)
    assert np.all(np.abs(fold - full) <= 1e-9 * scale + 1e-9 * np.abs(full).max())


@pytest.mark.parametrize("view", ["ap", "lateral"])
@pytest.mark.parametrize("pose", [(0.0, 0.0, 0.0, 0.0, 0.0, 0.0), (20.0, 0.0, 0.0, 1.0, 0.5, -0.5),
                                  (0.0, 90.0, 0.0, 0.0, 0.0, 0.0)])
def test_fold_matches_oracle_fold_convention_at_lattice_exact_poses(view, pose):
    """Where sample coordinates sit exactly on lattice planes the interpolant has a kink:
    the fold's one-sided derivatives are the oracle's ``fold`` convention (and the
    values and first four sums equal the plain full slice's)."""
    rng = np.random.default_rng(9)
    vol = np.zeros((6, 7, 8))
    vol[1:5, 2:6, 1:7] = rng.uniform(0.0, 1.0, (4, 4, 6))
    F = O.make_spectrum(vol)
    H, W = O.slice_dims(F.shape, view)
    T = rng.standard_normal((H, W)) + 1j * rng.standard_normal((H, W))
    pose = np.array(pose)
    fold = _moments_fold(F, pose, view, T)
    ref = _moments_full(F, pose, view, T, fold=True)
    np.testing.assert_allclose(fold, ref, atol=1e-9 * np.abs(ref).max(), rtol=1e-9)
    plain = _moments_full(F, pose, view, T)
    np.testing.assert_allclose(fold[:4], plain[:4], atol=1e-9 * np.abs(plain).max(), rtol=1e-9)
