"""FP64 NumPy/SciPy oracle for the parallel-beam Diff-FVR path (convention N).

TEST INFRASTRUCTURE ONLY -- imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by the product package.

Each function restates one piece of the reference, cited as file:line in
``src/fvr/`` of pareelamre/diff-fvr.  It makes the same library calls as the
reference where the reference calls a library (``scipy.fft``,
``scipy.ndimage.map_coordinates``).  The functions marked "analytic" are this
build's own exact pose adjoint (the reference computes rotation gradients by
central finite differences); the HIP kernel implements that adjoint.

How each function is pinned (see ``tests/test_oracle_n.py``):

* ``make_spectrum``            -- Fourier-slice theorem at the identity pose
                                  (AP slice == 2D FFT of the y-projection of the
                                  premultiplied padded volume); single-voxel
                                  delta gives an exact phase ramp; the sinc^2
                                  window against its closed form.
* ``euler_to_rotation`` /
  ``euler_to_basis``           -- orthonormality; 90-degree special cases.
* ``extract_frequency_slice``  -- the reference's known answer
                                  (``tests/test_physical_units.py:28-34``: a
                                  4/-2 mm shift at 2 mm spacing equals
                                  ``np.roll`` by (-1, 2) pixels, atol 1e-9).
* ``trilinear_value_and_grad`` -- value equals ``scipy.ndimage.map_coordinates``
                                  (order=1, mode='constant') on random and
                                  boundary coordinates; gradient equals a
                                  brute-force per-corner derivative and central
                                  differences away from cell faces.
* ``compute_spectral_ncc``     -- NCC(S,S)=1, scale invariance, zero norm -> 0,
                                  Parseval against the spatial correlation.
* ``ncc_and_grad_translation`` -- FP64 central differences of the exact loss.
* ``biplanar_loss_and_gradients`` (reference FD scheme) -- loss equals the sum
                                  of two ``compute_spectral_ncc`` terms; FD
                                  rotation gradients agree with the analytic
                                  adjoint to O(delta^2).
* ``biplanar_loss_and_grad_analytic`` -- FP64 central differences (h=1e-3 deg /
                                  voxel) and float64 torch autograd of an
                                  independent stencil restatement.
* ``adam_register``            -- the Adam closed form on a quadratic; with the
                                  multistart options (betas, gradient clamp,
                                  cosine factor, no pose clip) it equals
                                  ``torch.optim.Adam`` + ``LambdaLR``, the calls
                                  of ``experiments/exp30_gpu_multistart_registration.py:140-183``.
* ``compute_mtre_6dof``        -- the reference's closed forms 6.0, 3.0, 16.0
                                  (``tests/test_physical_units.py:52-59``).
* ``prepare_spectrum``         -- a target rendered at pose p matches the slice
                                  at pose p for the identity pose (Fourier-slice
                                  theorem with the crop/pad bookkeeping).
"""

from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy as np
from scipy import fft
from scipy.ndimage import map_coordinates

DEG = math.pi / 180.0


# ---------------------------------------------------------------------------
# a1: padded, premultiplied, centred 3D spectrum  (diff_fvr.py:22-54)
# ---------------------------------------------------------------------------
def sinc2_correction(n_orig: int, n_pad: int) -> np.ndarray:
    """1 / max(sinc^2((i - n/2)/p), 0.05)   (diff_fvr.py:35-39)."""
    coords = (np.arange(n_orig) - n_orig / 2.0) / n_pad
    s = np.sinc(coords)
    return 1.0 / np.maximum(s ** 2, 0.05)


def padded_shape(shape, pad_factor: float = 2.0) -> Tuple[int, int, int]:
    """int(round(n * pad_factor)) per axis (diff_fvr.py:30-32)."""
    return tuple(int(round(n * pad_factor)) for n in shape)


def make_padded_volume(volume: np.ndarray, pad_factor: float = 2.0,
                       premultiply: bool = True) -> np.ndarray:
    """Premultiplied volume zero-padded at offset (p - n)//2 (diff_fvr.py:27-51)."""
    vol = np.asarray(volume, dtype=np.float64)
    nz, ny, nx = vol.shape
    pz, py, px = padded_shape(vol.shape, pad_factor)
    if premultiply:
        vol = vol * (sinc2_correction(nz, pz)[:, None, None]
                     * sinc2_correction(ny, py)[None, :, None]
                     * sinc2_correction(nx, px)[None, None, :])
    padded = np.zeros((pz, py, px), dtype=np.float64)
    sz, sy, sx = (pz - nz) // 2, (py - ny) // 2, (px - nx) // 2
    padded[sz:sz + nz, sy:sy + ny, sx:sx + nx] = vol
    return padded


def make_spectrum(volume: np.ndarray, pad_factor: float = 2.0,
                  premultiply: bool = True) -> np.ndarray:
    """fftshift(fftn(ifftshift(padded)))  complex128, index (z, y, x) (diff_fvr.py:54)."""
    padded = make_padded_volume(volume, pad_factor, premultiply)
    return fft.fftshift(fft.fftn(fft.ifftshift(padded)))


# ---------------------------------------------------------------------------
# a2: pose -> slice basis (diff_fvr.py:56-104) and its derivative (analytic)
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


def _drz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[0.0, 0.0, 0.0], [0.0, -s, -c], [0.0, c, -s]])


def _dry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[-s, 0.0, c], [0.0, 0.0, 0.0], [-c, 0.0, -s]])


def _drx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[-s, -c, 0.0], [c, -s, 0.0], [0.0, 0.0, 0.0]])


def euler_to_rotation(yaw_deg, pitch_deg, roll_deg) -> np.ndarray:
    """rot = Rz(yaw) @ Ry(roll) @ Rx(pitch) in (Z, Y, X) index space (diff_fvr.py:64-87)."""
    return _rz(yaw_deg * DEG) @ _ry(roll_deg * DEG) @ _rx(pitch_deg * DEG)


VIEW_BASES = {
    # (u_base, v_base) in (wz, wy, wx)   (diff_fvr.py:90-100)
    "lateral": (np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0])),
    "ap": (np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])),
    "axial": (np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0])),
}


def _view_key(view: str) -> str:
    return view if view in ("lateral", "ap") else "axial"


def euler_to_basis(yaw_deg, pitch_deg, roll_deg, view="lateral"):
    """u = rot @ u_base, v = rot @ v_base  (diff_fvr.py:102-104)."""
    rot = euler_to_rotation(yaw_deg, pitch_deg, roll_deg)
    ub, vb = VIEW_BASES[_view_key(view)]
    return rot @ ub, rot @ vb


def euler_basis_jacobian(yaw_deg, pitch_deg, roll_deg, view="lateral"):
    """(du/dtheta, dv/dtheta) as (3, 3) arrays: column j = derivative w.r.t.
    pose[j] in DEGREES (j = yaw, pitch, roll).  Analytic; not in the reference."""
    y, p, r = yaw_deg * DEG, pitch_deg * DEG, roll_deg * DEG
    Rz, Ry, Rx = _rz(y), _ry(r), _rx(p)
    d_yaw = _drz(y) @ Ry @ Rx * DEG
    d_pitch = Rz @ Ry @ _drx(p) * DEG
    d_roll = Rz @ _dry(r) @ Rx * DEG
    ub, vb = VIEW_BASES[_view_key(view)]
    du = np.stack([d_yaw @ ub, d_pitch @ ub, d_roll @ ub], axis=1)
    dv = np.stack([d_yaw @ vb, d_pitch @ vb, d_roll @ vb], axis=1)
    return du, dv


# ---------------------------------------------------------------------------
# a3/a4: slice extraction (diff_fvr.py:106-150)
# ---------------------------------------------------------------------------
def slice_dims(pad_shape, view) -> Tuple[int, int]:
    """(slice_h, slice_w) per view (diff_fvr.py:121-126)."""
    pz, py, px = pad_shape
    if view == "lateral":
        return pz, py
    if view == "ap":
        return pz, px
    return py, px


def slice_grid(slice_h: int, slice_w: int):
    """uu, vv = meshgrid(linspace(-w/2, w/2, w, endpoint=False), ...) (diff_fvr.py:131-133)."""
    u = np.linspace(-slice_w / 2.0, slice_w / 2.0, slice_w, endpoint=False)
    v = np.linspace(-slice_h / 2.0, slice_h / 2.0, slice_h, endpoint=False)
    return np.meshgrid(u, v)


def slice_coords(pad_shape, u_vec, v_vec, view):
    """(3, K) sample coordinates (z, y, x) of one slice (diff_fvr.py:128-139)."""
    slice_h, slice_w = slice_dims(pad_shape, view)
    uu, vv = slice_grid(slice_h, slice_w)
    c = np.array(pad_shape, dtype=np.float64) / 2.0
    coords = np.array([(uu * u_vec[i] + vv * v_vec[i] + c[i]).ravel() for i in range(3)])
    return coords, uu, vv


def phase_ramp(uu, vv, tx, ty, slice_h, slice_w):
    """exp(-2 pi i (uu tx / W + vv ty / H))  (diff_fvr.py:145-148)."""
    return np.exp(-2j * np.pi * (uu * tx / slice_w + vv * ty / slice_h))


def extract_frequency_slice(f_3d, yaw_deg=0.0, pitch_deg=0.0, roll_deg=0.0,
                            tx=0.0, ty=0.0, view="lateral", interp_order=1):
    """Trilinear (scipy, order 1, mode='constant') slice + phase ramp (diff_fvr.py:106-150)."""
    pad_shape = f_3d.shape
    slice_h, slice_w = slice_dims(pad_shape, view)
    u_vec, v_vec = euler_to_basis(yaw_deg, pitch_deg, roll_deg, view)
    coords, uu, vv = slice_coords(pad_shape, u_vec, v_vec, view)
    re = map_coordinates(f_3d.real, coords, order=interp_order, mode="constant", cval=0.0)
    im = map_coordinates(f_3d.imag, coords, order=interp_order, mode="constant", cval=0.0)
    f_slice = (re + 1j * im).reshape(slice_h, slice_w)
    if tx != 0.0 or ty != 0.0:
        f_slice = f_slice * phase_ramp(uu, vv, tx, ty, slice_h, slice_w)
    return f_slice


def render_drr(f_3d, orig_shape, yaw_deg=0.0, pitch_deg=0.0, roll_deg=0.0, tx=0.0, ty=0.0,
               view="lateral", interp_order=1, crop_to_roi=True):
    """ifft2 of the slice, scale, ReLU, centre crop (diff_fvr.py:152-182)."""
    f_slice = extract_frequency_slice(f_3d, yaw_deg, pitch_deg, roll_deg, tx, ty, view, interp_order)
    slice_h, slice_w = f_slice.shape
    drr = np.real(fft.fftshift(fft.ifft2(fft.ifftshift(f_slice))))
    pz, py, px = f_3d.shape
    drr = np.maximum(drr * (pz * py * px) / (slice_h * slice_w), 0.0)
    if crop_to_roi:
        nz, ny, nx = orig_shape
        target_h = nz if view in ("lateral", "ap") else ny
        target_w = ny if view == "lateral" else nx
        sh, sw = (slice_h - target_h) // 2, (slice_w - target_w) // 2
        drr = drr[sh:sh + target_h, sw:sw + target_w]
    return drr


# ---------------------------------------------------------------------------
# a5/a6: Spectral-NCC, translation closed form, reference biplanar scheme
# ---------------------------------------------------------------------------
def compute_spectral_ncc(f_cand, f_targ) -> float:
    """Re sum(S conj T) / (||S|| ||T||), 0 if a norm is 0 (diff_fvr.py:184-195)."""
    inner = np.real(np.sum(f_cand * np.conj(f_targ)))
    ns, nt = np.linalg.norm(f_cand), np.linalg.norm(f_targ)
    if ns * nt == 0:
        return 0.0
    return float(inner / (ns * nt))


def ncc_and_grad_translation(f_cand, f_targ):
    """(ncc, dncc/du, dncc/dv) with the residual formulation (gradients.py:27-74)."""
    slice_h, slice_w = f_cand.shape
    ns, nt = np.linalg.norm(f_cand), np.linalg.norm(f_targ)
    if ns * nt == 0:
        return 0.0, 0.0, 0.0
    ncc = float(np.real(np.sum(f_cand * np.conj(f_targ))) / (ns * nt))
    r_res = (f_targ - ncc * (nt / ns) * f_cand) / (ns * nt)
    uu, vv = slice_grid(slice_h, slice_w)
    fu = -2j * np.pi * (uu / slice_w) * f_cand
    fv = -2j * np.pi * (vv / slice_h) * f_cand
    return (ncc, float(np.real(np.sum(fu * np.conj(r_res)))),
            float(np.real(np.sum(fv * np.conj(r_res)))))


def biplanar_loss_and_gradients(f_3d, pose, f_target_ap, f_target_lat, delta_rot_deg=0.2):
    """Reference scheme: closed-form translations, central-FD rotations (gradients.py:76-135).

    14 slice extractions: AP/LAT base plus +-delta per Euler angle.
    """
    yaw, pitch, roll, tx, ty, tz = [float(x) for x in pose]
    f_ap = extract_frequency_slice(f_3d, yaw, pitch, roll, tx=tx, ty=tz, view="ap")
    ncc_ap, d_ap_dtx, d_ap_dtz = ncc_and_grad_translation(f_ap, f_target_ap)
    f_lat = extract_frequency_slice(f_3d, yaw, pitch, roll, tx=ty, ty=tz, view="lateral")
    ncc_lat, d_lat_dty, d_lat_dtz = ncc_and_grad_translation(f_lat, f_target_lat)
    loss = 0.5 * (1.0 - ncc_ap) + 0.5 * (1.0 - ncc_lat)
    grad_rot = np.zeros(3)
    base = np.array([yaw, pitch, roll], dtype=np.float64)
    for i in range(3):
        pp, pm = base.copy(), base.copy()
        pp[i] += delta_rot_deg
        pm[i] -= delta_rot_deg
        ap_p = compute_spectral_ncc(extract_frequency_slice(f_3d, *pp, tx=tx, ty=tz, view="ap"), f_target_ap)
        ap_m = compute_spectral_ncc(extract_frequency_slice(f_3d, *pm, tx=tx, ty=tz, view="ap"), f_target_ap)
        la_p = compute_spectral_ncc(extract_frequency_slice(f_3d, *pp, tx=ty, ty=tz, view="lateral"), f_target_lat)
        la_m = compute_spectral_ncc(extract_frequency_slice(f_3d, *pm, tx=ty, ty=tz, view="lateral"), f_target_lat)
        grad_rot[i] = -0.5 * ((ap_p - ap_m) / (2.0 * delta_rot_deg) + (la_p - la_m) / (2.0 * delta_rot_deg))
    grad = np.array([grad_rot[0], grad_rot[1], grad_rot[2],
                     -0.5 * d_ap_dtx, -0.5 * d_lat_dty, -0.5 * (d_ap_dtz + d_lat_dtz)])
    return float(loss), grad


def biplanar_loss(f_3d, pose, f_target_ap, f_target_lat) -> float:
    """L = 1/2 (1 - NCC_AP) + 1/2 (1 - NCC_LAT) with AP (tx,tz), LAT (ty,tz) (gradients.py:95-102)."""
    yaw, pitch, roll, tx, ty, tz = [float(x) for x in pose]
    ap = compute_spectral_ncc(extract_frequency_slice(f_3d, yaw, pitch, roll, tx=tx, ty=tz, view="ap"), f_target_ap)
    la = compute_spectral_ncc(extract_frequency_slice(f_3d, yaw, pitch, roll, tx=ty, ty=tz, view="lateral"), f_target_lat)
    return 0.5 * (1.0 - ap) + 0.5 * (1.0 - la)


# ---------------------------------------------------------------------------
# analytic: trilinear value + spatial gradient with scipy's constant-mode rule
# ---------------------------------------------------------------------------
def trilinear_value_and_grad(f_3d, coords, left=None, wrap=False):
    """Value and d/d(z, y, x) of scipy's order-1 'constant' interpolant.

    scipy rule (pinned by brute force in the tests): a sample whose coordinate
    lies outside [0, n-1] on any axis is 0; otherwise the usual trilinear blend
    of the corners floor(p) and floor(p)+1.  Derivative convention (this
    build's): forward difference across the cell [floor(p), floor(p)+1], with
    a corner index of n (reached only when p == n-1 exactly) reading 0.
    Returns I (K,) complex and dI (3, K) complex.

    The two options matter only where a coordinate is exactly an integer (the
    interpolant has a kink there; the value is unaffected):
    ``left`` (bool (K,)) takes the backward difference, cell [ceil(p)-1, ceil(p)],
    for those samples; ``wrap`` reads corner index n as index 0 (the periodic
    spectrum).  Together they give the one-sided derivatives of the Hermitian-fold
    kernel: forward on the wrapped lattice for the samples it evaluates,
    backward for the mirrors it infers from them.
    """
    shp = np.array(f_3d.shape)
    p = np.asarray(coords, dtype=np.float64)
    inb = np.all((p >= 0.0) & (p <= (shp - 1)[:, None]), axis=0)
    i0 = np.floor(p).astype(np.int64)
    if left is not None:
        on = (p == np.floor(p)) & np.asarray(left, bool)[None, :]
        i0 = np.where(on, i0 - 1, i0)
    t = p - i0
    K = p.shape[1]
    vals = {}
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                iz, iy, ix = i0[0] + dz, i0[1] + dy, i0[2] + dx
                if wrap:
                    iz, iy, ix = (np.where(q == n, 0, q) for q, n in zip((iz, iy, ix), shp))
                ok = inb & (iz >= 0) & (iz < shp[0]) & (iy >= 0) & (iy < shp[1]) & (ix >= 0) & (ix < shp[2])
                v = np.zeros(K, dtype=np.complex128)
                v[ok] = f_3d[iz[ok], iy[ok], ix[ok]]
                vals[(dz, dy, dx)] = v
    tz, ty, tx = t
    w = {0: (1 - tz, 1 - ty, 1 - tx), 1: (tz, ty, tx)}
    I = np.zeros(K, dtype=np.complex128)
    dI = np.zeros((3, K), dtype=np.complex128)
    for (dz, dy, dx), v in vals.items():
        wz = tz if dz else 1 - tz
        wy = ty if dy else 1 - ty
        wx = tx if dx else 1 - tx
        sz = 1.0 if dz else -1.0
        sy = 1.0 if dy else -1.0
        sx = 1.0 if dx else -1.0
        I += wz * wy * wx * v
        dI[0] += sz * wy * wx * v
        dI[1] += wz * sy * wx * v
        dI[2] += wz * wy * sx * v
    I[~inb] = 0.0
    dI[:, ~inb] = 0.0
    return I, dI


def fold_mirror_mask(slice_h: int, slice_w: int) -> np.ndarray:
    """(h*w,) True for the samples the Hermitian-fold kernel infers as mirrors:
    rows 1 .. h/2-1 (columns >= 1) and row h/2 left of the centre (columns >= 1)."""
    i, j = np.meshgrid(np.arange(slice_h), np.arange(slice_w), indexing="ij")
    hh, hw = slice_h // 2, slice_w // 2
    return (((i >= 1) & (i < hh) & (j >= 1)) | ((i == hh) & (j >= 1) & (j < hw))).ravel()


def view_ncc_and_grad_analytic(f_3d, yaw, pitch, roll, a, b, view, f_targ, fold=False):
    """Exact NCC and its derivative w.r.t. (yaw, pitch, roll, a, b) for one view.

    S = I(c + uu u + vv v) * exp(-2 pi i (uu a/W + vv b/H));
    dNCC = Re<dS,T>/(|S||T|) - NCC Re<dS,S>/|S|^2.
    ``fold``: the one-sided derivative convention of the Hermitian-fold kernel at
    lattice-exact coordinates (see ``trilinear_value_and_grad``).
    Returns (ncc, d (5,), aux) where aux holds the raw sums.
    """
    pad_shape = f_3d.shape
    slice_h, slice_w = slice_dims(pad_shape, view)
    u_vec, v_vec = euler_to_basis(yaw, pitch, roll, view)
    du, dv = euler_basis_jacobian(yaw, pitch, roll, view)
    coords, uu, vv = slice_coords(pad_shape, u_vec, v_vec, view)
    if fold:
        I, dI = trilinear_value_and_grad(f_3d, coords, left=fold_mirror_mask(slice_h, slice_w), wrap=True)
    else:
        I, dI = trilinear_value_and_grad(f_3d, coords)
    uu, vv = uu.ravel(), vv.ravel()
    ph = np.exp(-2j * np.pi * (uu * a / slice_w + vv * b / slice_h))
    S = I * ph
    T = np.asarray(f_targ).ravel()
    ns, nt = np.linalg.norm(S), np.linalg.norm(T)
    if ns * nt == 0:
        return 0.0, np.zeros(5), {}
    A = float(np.real(np.sum(S * np.conj(T))))
    ncc = A / (ns * nt)
    d = np.zeros(5)
    for j in range(3):
        dp = uu[None, :] * du[:, j][:, None] + vv[None, :] * dv[:, j][:, None]
        dS = ph * np.sum(dI * dp, axis=0)
        G = float(np.real(np.sum(dS * np.conj(T))))
        H = float(np.real(np.sum(dS * np.conj(S))))
        d[j] = G / (ns * nt) - ncc * H / ns ** 2
    dSa = -2j * np.pi * (uu / slice_w) * S
    dSb = -2j * np.pi * (vv / slice_h) * S
    d[3] = float(np.real(np.sum(dSa * np.conj(T)))) / (ns * nt) - ncc * float(np.real(np.sum(dSa * np.conj(S)))) / ns ** 2
    d[4] = float(np.real(np.sum(dSb * np.conj(T)))) / (ns * nt) - ncc * float(np.real(np.sum(dSb * np.conj(S)))) / ns ** 2
    return ncc, d, {"A": A, "E": ns ** 2, "T2": nt ** 2, "S": S.reshape(slice_h, slice_w)}


def biplanar_loss_and_grad_analytic(f_3d, pose, f_target_ap, f_target_lat, fold=False):
    """Exact loss and 6-DoF gradient of L = 1/2(1-NCC_AP) + 1/2(1-NCC_LAT).

    Same loss as gradients.py:95-102; rotation derivatives exact instead of the
    reference's central differences (gradients.py:112-132).  ``fold`` selects the
    Hermitian-fold kernel's one-sided derivatives at lattice-exact poses.
    """
    yaw, pitch, roll, tx, ty, tz = [float(x) for x in pose]
    n_ap, d_ap, _ = view_ncc_and_grad_analytic(f_3d, yaw, pitch, roll, tx, tz, "ap", f_target_ap, fold)
    n_la, d_la, _ = view_ncc_and_grad_analytic(f_3d, yaw, pitch, roll, ty, tz, "lateral", f_target_lat, fold)
    loss = 0.5 * (1.0 - n_ap) + 0.5 * (1.0 - n_la)
    grad = np.zeros(6)
    grad[:3] = -0.5 * (d_ap[:3] + d_la[:3])
    grad[3] = -0.5 * d_ap[3]
    grad[4] = -0.5 * d_la[3]
    grad[5] = -0.5 * (d_ap[4] + d_la[4])
    return float(loss), grad


# ---------------------------------------------------------------------------
# a7: target spectrum (core_engine.py:71-85)
# ---------------------------------------------------------------------------
def prepare_spectrum(img, slice_h, slice_w):
    """Zero-pad centred into (slice_h, slice_w), fftshift(fft2(ifftshift(.)))."""
    th, tw = img.shape
    padded = np.zeros((slice_h, slice_w), dtype=np.float64)
    sh, sw = (slice_h - th) // 2, (slice_w - tw) // 2
    padded[sh:sh + th, sw:sw + tw] = img
    return fft.fftshift(fft.fft2(fft.ifftshift(padded)))


# ---------------------------------------------------------------------------
# a8: NumPy Adam (gradient_optimizer.py:24-90)
# ---------------------------------------------------------------------------
def adam_register(loss_grad_fn, init_pose, lr_rot=0.25, lr_trans=0.35, max_iter=40, tol=1e-4, beta1=0.9,
                  beta2=0.999, eps=1e-8, limits=(20.0, 25.0), gclip=None, cosine_t=0) -> Dict:
    """Adam with decoupled lr, clipping to +-20 deg / +-25, early stop on max|step| < tol
    (gradient_optimizer.py:24-90).  Multistart options (exp30:140-183): ``gclip`` =
    (rot, trans) gradient clamp, ``cosine_t`` > 0 scales lr by 0.1 + 0.9 (1 + cos(pi s/T))/2
    at scheduler step s = min(step - 1, T); ``limits=None`` leaves the pose unclipped."""
    pose = np.array(init_pose, dtype=np.float64)
    m = np.zeros(6)
    v = np.zeros(6)
    b1, b2 = beta1, beta2
    lr = np.array([lr_rot] * 3 + [lr_trans] * 3)
    lim = None if limits is None else np.array([limits[0]] * 3 + [limits[1]] * 3)
    gc = None if gclip is None else np.array([gclip[0]] * 3 + [gclip[1]] * 3)
    best_loss, best_pose = float("inf"), pose.copy()
    traj = []
    step = 0
    for step in range(1, max_iter + 1):
        loss, grad = loss_grad_fn(pose)
        traj.append((step, loss, pose.copy()))
        if loss < best_loss:
            best_loss, best_pose = loss, pose.copy()
        if gc is not None:
            grad = np.clip(grad, -gc, gc)
        m = b1 * m + (1 - b1) * grad
        v = b2 * v + (1 - b2) * grad ** 2
        f = 1.0
        if cosine_t > 0:
            f = 0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * min(step - 1, cosine_t) / cosine_t))
        upd = f * lr * ((m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps))
        pose = pose - upd
        if lim is not None:
            pose = np.clip(pose, -lim, lim)
        if np.max(np.abs(upd)) < tol:
            break
    return {"success": bool(best_loss < 0.05), "final_pose": best_pose, "final_loss": float(best_loss),
            "iterations": step, "evaluations": step * 7, "trajectory": traj}


# ---------------------------------------------------------------------------
# a9: mTRE (registration_3d2d.py:112-175)
# ---------------------------------------------------------------------------
def compute_mtre_6dof(true_pose, est_pose, volume_shape, spacing=(1.0, 1.0, 1.0), translation_units="voxels"):
    spacing = np.asarray(spacing, dtype=float)
    nz, ny, nx = volume_shape
    c = np.array([nz / 2.0, ny / 2.0, nx / 2.0])
    r = min(nz, ny, nx) * 0.25
    corners = np.array([[c[0] + sz * r, c[1] + sy * r, c[2] + sx * r]
                        for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)])

    def tf(pts, pose):
        yaw, pitch, roll, tx, ty, tz = pose
        rot = euler_to_rotation(yaw, pitch, roll)
        tr = np.array([tz, ty, tx], dtype=float)
        if translation_units == "voxels":
            tr = tr * spacing
        return ((pts - c) * spacing) @ rot.T + tr

    return float(np.mean(np.linalg.norm(tf(corners, true_pose) - tf(corners, est_pose), axis=1)))
