"""FP64 NumPy oracle for the reference's torch path (convention T) and its slice operator.

TEST INFRASTRUCTURE ONLY -- imported by ``tests/`` (and nothing in the product
package).  Each function restates one piece of ``src/fvr/torch_engine.py`` or
``src/fvr/triton_slice.py`` of pareelamre/diff-fvr (cited per function) in
float64, with the reference's float32 storage replaced by float64 so that the
HIP results are compared against the exact values.  The ``*_grad`` functions are
this build's exact adjoints (the reference obtains them from torch autograd).

How each function is pinned (``tests/test_oracle_t.py``):

* ``t_spectrum``               -- ``numpy.fft`` of the centred zero-padded
                                  volume, checked against the Fourier-slice
                                  theorem (the z = centre plane of the spectrum is
                                  the 2D FFT of the z-projection) and against a
                                  single-voxel delta's closed-form phase.
* ``t_rotation``               -- orthonormality, 90-degree special cases, and
                                  ``t_rotation_derivs`` by central differences.
* ``slice_per_corner`` (+grad) -- value equals ``torch.nn.functional.grid_sample``
                                  (float64, align_corners=True, zeros), the
                                  library routine the reference's own test
                                  compares its stencil with
                                  (``tests/test_cone_beam_fvr.py:152-205``);
                                  gradient equals float64 torch autograd of that
                                  ``grid_sample`` call.
* ``fourier_slice_3d`` / ``fourier_slice_3d_grad_uv``
                               -- as above, through the (B_vol, 2, D, H, W)
                                  planar layout of ``triton_slice.py:648-661``.
* ``hp_filter``                -- closed form at the centre (0 at r=0 when the
                                  grid hits 0) and its limit 1 far away.
* ``spectral_ncc_core`` (+grad) -- loss in [0, 2]; NCC of a slice with its own
                                  high-passed copy is 1 (up to 1e-8 eps);
                                  gradient by FP64 central differences
                                  (h = 1e-3 deg / voxel).
* ``spectral_ncc_loss`` (+grad) -- image-space loss: without the high-pass
                                  weight, Parseval (NCC = sum p t / (|p| |t|) in
                                  the image domain); with it, the same identity
                                  for the high-passed images formed by an inverse
                                  FFT; NCC(p, p) = 1, NCC(p, -p) = -1, scale
                                  invariance, batch mean; gradient w.r.t. the
                                  predicted image by FP64 central differences.
* ``forward_drr``              -- the inverse FFT of the slice at the identity
                                  pose equals the ray sum along z of the padded
                                  volume sampled on the detector grid when the
                                  grid hits lattice points (out_res = 2N - 1).
"""

from __future__ import annotations

import math
from typing import Tuple

import numpy as np

DEG = math.pi / 180.0


# ---------------------------------------------------------------------------
# spectrum (torch_engine.py:55-88): centred zero pad, fftshift(fftn(.)), no
# ifftshift, no sinc^2
# ---------------------------------------------------------------------------
def t_padded_volume(volume: np.ndarray, pad_factor: float = 2.0) -> np.ndarray:
    d, h, w = volume.shape
    pd, ph, pw = (int(round(n * pad_factor)) for n in (d, h, w))
    out = np.zeros((pd, ph, pw), dtype=np.float64)
    out[(pd - d) // 2:(pd - d) // 2 + d, (ph - h) // 2:(ph - h) // 2 + h, (pw - w) // 2:(pw - w) // 2 + w] = volume
    return out


def t_spectrum(volume: np.ndarray, pad_factor: float = 2.0) -> np.ndarray:
    """Complex128 (D, H, W) spectrum of convention T (torch_engine.py:82-88)."""
    return np.fft.fftshift(np.fft.fftn(t_padded_volume(np.asarray(volume, dtype=np.float64), pad_factor)))


# ---------------------------------------------------------------------------
# rotation and view axes (torch_engine.py:138-211)
# ---------------------------------------------------------------------------
def _mrz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _mry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def _mrx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def t_rotation(yaw_deg, pitch_deg, roll_deg) -> np.ndarray:
    """R = Rz(yaw) Ry(pitch) Rx(roll), standard xyz axes (torch_engine.py:138-174)."""
    return _mrz(yaw_deg * DEG) @ _mry(pitch_deg * DEG) @ _mrx(roll_deg * DEG)


def t_rotation_derivs(yaw_deg, pitch_deg, roll_deg) -> np.ndarray:
    """dR/d(yaw, pitch, roll) per degree, (3, 3, 3)."""
    a, b, c = yaw_deg * DEG, pitch_deg * DEG, roll_deg * DEG

    # derivatives of the single-axis rotations: d/da [[c,-s],[s,c]] = [[-s,-c],[c,-s]]
    def drz(x):
        cc, s = math.cos(x), math.sin(x)
        return np.array([[-s, -cc, 0.0], [cc, -s, 0.0], [0.0, 0.0, 0.0]])

    def dry(x):
        cc, s = math.cos(x), math.sin(x)
        return np.array([[-s, 0.0, cc], [0.0, 0.0, 0.0], [-cc, 0.0, -s]])

    def drx(x):
        cc, s = math.cos(x), math.sin(x)
        return np.array([[0.0, 0.0, 0.0], [0.0, -s, -cc], [0.0, cc, -s]])

    return np.stack([drz(a) @ _mry(b) @ _mrx(c), _mrz(a) @ dry(b) @ _mrx(c), _mrz(a) @ _mry(b) @ drx(c)]) * DEG


RZ90 = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])


def t_center(pad_shape) -> np.ndarray:
    """(cx, cy, cz) = ((W-1)/2, (H-1)/2, (D-1)/2) (torch_engine.py:200-205)."""
    D, H, W = pad_shape
    return np.array([0.5 * max(W - 1, 1), 0.5 * max(H - 1, 1), 0.5 * max(D - 1, 1)])


def t_view_axes(pose, view, pad_shape):
    """u_xyz, v_xyz and du/dtheta, dv/dtheta (3 angles x 3) of one pose and view
    (torch_engine.py:188-211; lateral composes R_pose @ Rz(90 deg))."""
    R = t_rotation(*pose[:3])
    dR = t_rotation_derivs(*pose[:3])
    if view == "lateral":
        R = R @ RZ90
        dR = dR @ RZ90
    s = t_center(pad_shape)
    return R[0] * s, R[1] * s, dR[:, 0, :] * s, dR[:, 1, :] * s


def t_grid(out_res: int) -> np.ndarray:
    """linspace(-1, 1, out_res) (torch_engine.py:115-118)."""
    return np.linspace(-1.0, 1.0, out_res)


# ---------------------------------------------------------------------------
# per-corner trilinear sampling (triton_slice.py:664-715, grid_sample
# align_corners=True, padding 'zeros')
# ---------------------------------------------------------------------------
def slice_per_corner(vol: np.ndarray, p_xyz: np.ndarray, with_grad: bool = False):
    """Trilinear value of ``vol`` (D, H, W) at voxel coordinates ``p_xyz`` (..., 3) with
    every out-of-range corner read as 0; optionally d/dp (..., 3) in xyz order."""
    p = np.asarray(p_xyz, dtype=np.float64)
    D, H, W = vol.shape
    x0 = np.floor(p[..., 0])
    y0 = np.floor(p[..., 1])
    z0 = np.floor(p[..., 2])
    tx, ty, tz = p[..., 0] - x0, p[..., 1] - y0, p[..., 2] - z0
    ix, iy, iz = x0.astype(np.int64), y0.astype(np.int64), z0.astype(np.int64)

    def g(dx, dy, dz):
        x, y, z = ix + dx, iy + dy, iz + dz
        ok = (x >= 0) & (x < W) & (y >= 0) & (y < H) & (z >= 0) & (z < D)
        return np.where(ok, vol[np.clip(z, 0, D - 1), np.clip(y, 0, H - 1), np.clip(x, 0, W - 1)], 0.0)

    c = {(dx, dy, dz): g(dx, dy, dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)}
    ox, oy, oz = 1.0 - tx, 1.0 - ty, 1.0 - tz
    val = 0.0
    for (dx, dy, dz), cv in c.items():
        val = val + cv * (tx if dx else ox) * (ty if dy else oy) * (tz if dz else oz)
    if not with_grad:
        return val
    gx = gy = gz = 0.0
    for (dx, dy, dz), cv in c.items():
        wx, wy, wz = (tx if dx else ox), (ty if dy else oy), (tz if dz else oz)
        sx, sy, sz = (1.0 if dx else -1.0), (1.0 if dy else -1.0), (1.0 if dz else -1.0)
        gx = gx + cv * sx * wy * wz
        gy = gy + cv * wx * sy * wz
        gz = gz + cv * wx * wy * sz
    return val, np.stack([gx, gy, gz], axis=-1)


def fourier_slice_3d(vol_ri: np.ndarray, u_xyz, v_xyz, fx, fy, center_xyz) -> np.ndarray:
    """(B_vol, N, 2, K) planar slices of a (B_vol, 2, D, H, W) volume (triton_slice.py:47-143, 648-661):
    p = center + fx u + fy v, per-corner zero padding."""
    u = np.asarray(u_xyz, np.float64).reshape(-1, 3)
    v = np.asarray(v_xyz, np.float64).reshape(-1, 3)
    fx = np.asarray(fx, np.float64).reshape(-1)
    fy = np.asarray(fy, np.float64).reshape(-1)
    p = np.asarray(center_xyz, np.float64)[None, None, :] + fx[None, :, None] * u[:, None, :] \
        + fy[None, :, None] * v[:, None, :]
    B = vol_ri.shape[0]
    out = np.empty((B, len(u), 2, len(fx)))
    for b in range(B):
        for ch in range(2):
            out[b, :, ch, :] = slice_per_corner(np.asarray(vol_ri[b, ch], np.float64), p)
    return out


def fourier_slice_3d_grad_uv(vol_ri, u_xyz, v_xyz, fx, fy, center_xyz, grad_out):
    """d(sum grad_out * slice)/du, /dv (N, 3) each (triton_slice.py:146-271, 306-339)."""
    u = np.asarray(u_xyz, np.float64).reshape(-1, 3)
    v = np.asarray(v_xyz, np.float64).reshape(-1, 3)
    fx = np.asarray(fx, np.float64).reshape(-1)
    fy = np.asarray(fy, np.float64).reshape(-1)
    p = np.asarray(center_xyz, np.float64)[None, None, :] + fx[None, :, None] * u[:, None, :] \
        + fy[None, :, None] * v[:, None, :]
    gp = np.zeros(p.shape)
    for b in range(vol_ri.shape[0]):
        for ch in range(2):
            _, dp = slice_per_corner(np.asarray(vol_ri[b, ch], np.float64), p, with_grad=True)
            gp += np.asarray(grad_out[b, :, ch, :], np.float64)[..., None] * dp
    gu = np.einsum("nk,nkc->nc", np.broadcast_to(fx[None, :], gp.shape[:2]), gp)
    gv = np.einsum("nk,nkc->nc", np.broadcast_to(fy[None, :], gp.shape[:2]), gp)
    return gu, gv


# ---------------------------------------------------------------------------
# slice spectrum, high-pass filter, Spectral-NCC (torch_engine.py:123-136,
# 176-234, 255-290)
# ---------------------------------------------------------------------------
def hp_filter(H: int, W: int, cutoff_sigma: float = 4.0) -> np.ndarray:
    """1 - exp(-r^2 / (2 sigma^2)) on linspace(-H/2, H/2, H) x linspace(-W/2, W/2, W) (:123-136)."""
    y = np.linspace(-H / 2, H / 2, H)
    x = np.linspace(-W / 2, W / 2, W)
    return 1.0 - np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / (2.0 * cutoff_sigma ** 2))


def _slice_points(pose, view, pad_shape, out_res):
    u, v, du, dv = t_view_axes(pose, view, pad_shape)
    f = t_grid(out_res)
    c = t_center(pad_shape)
    # rows follow fx (u), columns follow fy (v): meshgrid(.., indexing='ij') of torch_engine.py:116
    p = c[None, None, :] + f[:, None, None] * u[None, None, :] + f[None, :, None] * v[None, None, :]
    return p, f, u, v, du, dv


def extract_slice_spectrum(spec: np.ndarray, pose, view="ap", out_res=128, spacing0=1.0) -> np.ndarray:
    """Phase-modulated slice (out_res, out_res) complex (torch_engine.py:176-234)."""
    p, f, *_ = _slice_points(np.asarray(pose, np.float64), view, spec.shape, out_res)
    s = slice_per_corner(spec.real, p) + 1j * slice_per_corner(spec.imag, p)
    span = out_res * spacing0 + 1e-6
    ang = -2.0 * math.pi * (f[:, None] * pose[3] + f[None, :] * pose[4]) / span
    return s * np.exp(1j * ang)


def forward_drr(spec, pose, view="ap", out_res=128, spacing0=1.0) -> np.ndarray:
    """relu(real(ifft2(ifftshift(slice)))) (torch_engine.py:236-241)."""
    s = extract_slice_spectrum(spec, pose, view, out_res, spacing0)
    return np.maximum(np.real(np.fft.ifft2(np.fft.ifftshift(s))), 0.0)


def prepare_spectral_target(target_img: np.ndarray, high_pass=True, cutoff_sigma=4.0) -> np.ndarray:
    """fftshift(fft2(img)) [* hp] (torch_engine.py:255-267)."""
    H, W = target_img.shape[-2:]
    f = np.fft.fftshift(np.fft.fft2(np.asarray(target_img, np.float64)), axes=(-2, -1))
    return f * hp_filter(H, W, cutoff_sigma) if high_pass else f


def spectral_ncc_core(spec, poses, f_targ, view="ap", high_pass=True, cutoff_sigma=4.0, spacing0=1.0) -> float:
    """1 - mean_b NCC_b, NCC = sum Re(Sp T*) / (sqrt(E_p E_t) + 1e-8) (torch_engine.py:269-290)."""
    return spectral_ncc_core_and_grad(spec, poses, f_targ, view, high_pass, cutoff_sigma, spacing0,
                                      with_grad=False)[0]


def spectral_ncc_core_and_grad(spec, poses, f_targ, view="ap", high_pass=True, cutoff_sigma=4.0, spacing0=1.0,
                               with_grad=True):
    """Loss and its exact gradient w.r.t. the (B, 6) poses (tz has no effect: 0)."""
    poses = np.asarray(poses, np.float64).reshape(-1, 6)
    B = len(poses)
    out_res = int(np.asarray(f_targ).shape[-1])
    T = np.asarray(f_targ).reshape(-1, out_res, out_res)
    hp = hp_filter(out_res, out_res, cutoff_sigma) if high_pass else np.ones((out_res, out_res))
    span = out_res * spacing0 + 1e-6
    nccs = np.zeros(B)
    grad = np.zeros((B, 6))
    for b in range(B):
        pose = poses[b]
        Tb = T[b if len(T) > 1 else 0]
        p, f, u, v, du, dv = _slice_points(pose, view, spec.shape, out_res)
        if with_grad:
            Ir, gr = slice_per_corner(spec.real, p, with_grad=True)
            Ii, gi = slice_per_corner(spec.imag, p, with_grad=True)
        else:
            Ir, Ii = slice_per_corner(spec.real, p), slice_per_corner(spec.imag, p)
        ang = -2.0 * math.pi * (f[:, None] * pose[3] + f[None, :] * pose[4]) / span
        ph = np.exp(1j * ang)
        S = (Ir + 1j * Ii) * ph * hp
        A = np.sum(np.real(S * np.conj(Tb)))
        E = np.sum(np.abs(S) ** 2)
        Et = np.sum(np.abs(Tb) ** 2)
        D = math.sqrt(E * Et) + 1e-8
        nccs[b] = A / D
        if not with_grad:
            continue
        # dS/dp (xyz) per sample, then dp/dtheta = fx_i du + fy_j dv
        dS = (gr + 1j * gi) * (ph * hp)[..., None]
        fi = np.broadcast_to(f[:, None], S.shape)
        fj = np.broadcast_to(f[None, :], S.shape)
        for k in range(3):
            dp = fi[..., None] * du[k][None, None, :] + fj[..., None] * dv[k][None, None, :]
            dSk = np.sum(dS * dp, axis=-1)
            dA = np.sum(np.real(dSk * np.conj(Tb)))
            dE = 2.0 * np.sum(np.real(dSk * np.conj(S)))
            dD = math.sqrt(Et) * dE / (2.0 * math.sqrt(E)) if E > 0 else 0.0
            grad[b, k] = dA / D - A * dD / D ** 2
        # translations: dS/dtx = -2 pi i fx_i / span S, dS/dty = -2 pi i fy_j / span S
        grad[b, 3] = np.sum(np.real(-2j * math.pi * fi / span * S * np.conj(Tb))) / D
        grad[b, 4] = np.sum(np.real(-2j * math.pi * fj / span * S * np.conj(Tb))) / D
    loss = 1.0 - float(np.mean(nccs))
    return loss, -grad / B, nccs


# ---------------------------------------------------------------------------
# image-space Spectral-NCC loss (torch_engine.py:309-337)
# ---------------------------------------------------------------------------
def spectral_ncc_loss(pred, target, high_pass=True, cutoff_sigma=4.0, with_grad=False):
    """1 - mean_b NCC_b of (B, 1, H, W) images, NCC = sum Re(Fp Ft*) / (sqrt(Ep Et) + 1e-8) with
    F = fftshift(fft2(img)) [* hp] (torch_engine.py:309-337).  ``with_grad``: also d loss /
    d pred (B, 1, H, W): dA/dp = H W Re(ifft2(ifftshift(hp Ft))), dEp/dp = 2 H W
    Re(ifft2(ifftshift(hp Fp))) (F real-linear in p, fft2's adjoint is H W ifft2)."""
    p = np.asarray(pred, np.float64)
    t = np.asarray(target, np.float64)
    H, W = p.shape[-2:]
    hp = hp_filter(H, W, cutoff_sigma) if high_pass else np.ones((H, W))
    Fp = np.fft.fftshift(np.fft.fft2(p), axes=(-2, -1)) * hp
    Ft = np.fft.fftshift(np.fft.fft2(np.broadcast_to(t, p.shape)), axes=(-2, -1)) * hp
    A = np.sum(np.real(Fp * np.conj(Ft)), axis=(-2, -1))
    Ep = np.sum(np.abs(Fp) ** 2, axis=(-2, -1))
    Et = np.sum(np.abs(Ft) ** 2, axis=(-2, -1))
    D = np.sqrt(Ep * Et) + 1e-8
    ncc = A / D
    loss = 1.0 - float(np.mean(ncc))
    if not with_grad:
        return loss
    adj = lambda X: H * W * np.real(np.fft.ifft2(np.fft.ifftshift(X, axes=(-2, -1))))
    dA = adj(hp * Ft)
    dEp = 2.0 * adj(hp * Fp)
    sEp = np.sqrt(Ep)
    dD = np.where(sEp[..., None, None] > 0, np.sqrt(Et)[..., None, None] * dEp / (2.0 * np.where(sEp > 0, sEp, 1.0))
                  [..., None, None], 0.0)
    dncc = dA / D[..., None, None] - (A / D ** 2)[..., None, None] * dD
    return loss, -dncc / ncc.size
