"""Pins of the convention-T oracle (oracle/torch_engine_np.py) and of the host side of
the fused T path (item geometry, Jacobians, moment combine).  CPU only."""

import math

import numpy as np
import pytest
import torch

from oracle import torch_engine_np as T


def _vol(n=(10, 12, 14), seed=0):
    rng = np.random.default_rng(seed)
    v = np.zeros(n)
    c = np.array(n) / 2.0
    zz, yy, xx = np.meshgrid(*[np.arange(k) for k in n], indexing="ij")
    for _ in range(4):
        ax = rng.uniform(1.5, 3.5, 3)
        ctr = c + rng.uniform(-2, 2, 3)
        v += rng.uniform(0.01, 0.05) * ((((zz - ctr[0]) / ax[0]) ** 2 + ((yy - ctr[1]) / ax[1]) ** 2
                                         + ((xx - ctr[2]) / ax[2]) ** 2) <= 1.0)
    return v


def test_spectrum_fourier_slice_theorem_and_delta():
    v = _vol()
    S = T.t_spectrum(v)
    D, H, W = S.shape
    # the kz = 0 plane (index D//2 after fftshift) is the 2D FFT of the z-projection
    proj = T.t_padded_volume(v).sum(axis=0)
    np.testing.assert_allclose(S[D // 2], np.fft.fftshift(np.fft.fft2(proj)), atol=1e-11)
    # a delta at (z0, y0, x0) of the padded grid has spectrum exp(-2 pi i k.r) (k centred)
    d = np.zeros((5, 6, 7))
    d[2, 1, 3] = 1.0
    Sd = T.t_spectrum(d, pad_factor=2.0)
    P = Sd.shape
    pos = [(P[a] - d.shape[a]) // 2 + (2, 1, 3)[a] for a in range(3)]
    k = [np.fft.fftshift(np.fft.fftfreq(P[a])) for a in range(3)]
    ref = np.exp(-2j * np.pi * (k[0][:, None, None] * pos[0] + k[1][None, :, None] * pos[1]
                                + k[2][None, None, :] * pos[2]))
    np.testing.assert_allclose(Sd, ref, atol=1e-12)


def test_rotation_convention_and_derivatives():
    R = T.t_rotation(90.0, 0.0, 0.0)
    np.testing.assert_allclose(R @ [1.0, 0, 0], [0, 1.0, 0], atol=1e-15)  # yaw about z: x -> y
    R = T.t_rotation(0.0, 90.0, 0.0)
    np.testing.assert_allclose(R @ [0, 0, 1.0], [1.0, 0, 0], atol=1e-15)  # pitch about y: z -> x
    R = T.t_rotation(0.0, 0.0, 90.0)
    np.testing.assert_allclose(R @ [0, 1.0, 0], [0, 0, 1.0], atol=1e-15)  # roll about x: y -> z
    a = (12.0, -23.0, 31.0)
    R = T.t_rotation(*a)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-15)
    dR = T.t_rotation_derivs(*a)
    h = 1e-5
    for j in range(3):
        ap, am = list(a), list(a)
        ap[j] += h
        am[j] -= h
        fd = (T.t_rotation(*ap) - T.t_rotation(*am)) / (2 * h)
        np.testing.assert_allclose(dR[j], fd, atol=1e-10)


def _grid_sample_ref(vol, p, need_grad=False):
    """torch.nn.functional.grid_sample (float64, align_corners=True, zeros) at voxel coordinates p (K, 3) xyz."""
    D, H, W = vol.shape
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=need_grad)
    scale = torch.tensor([2.0 / (W - 1), 2.0 / (H - 1), 2.0 / (D - 1)], dtype=torch.float64)
    g = pt * scale - 1.0
    out = torch.nn.functional.grid_sample(torch.tensor(vol)[None, None], g.view(1, -1, 1, 1, 3), mode="bilinear",
                                          padding_mode="zeros", align_corners=True).reshape(-1)
    if not need_grad:
        return out.detach().numpy()
    w = torch.tensor(np.random.default_rng(5).standard_normal(out.shape[0]))
    (out * w).sum().backward()
    return out.detach().numpy(), pt.grad.numpy(), w.numpy()


def test_per_corner_stencil_equals_grid_sample_value_and_gradient():
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((6, 7, 8))
    D, H, W = vol.shape
    n = np.array([W, H, D], dtype=np.float64)
    p = rng.uniform(0, 1, (400, 3)) * (n + 1.2) - 1.1  # covers [-1.1, n + 0.1)
    # boundary coordinates: (-1, 0), n-1 exactly, (n-1, n), and just outside
    edge = []
    for a in range(3):
        for val in (-0.999, -0.5, -0.001, n[a] - 1.0, n[a] - 0.75, n[a] - 0.001, n[a] + 0.2, -1.2):
            q = np.array([W / 2.3, H / 2.1, D / 2.7])
            q[a] = val
            edge.append(q)
    p = np.concatenate([p, np.array(edge)])
    ref, gref, w = _grid_sample_ref(vol, p, need_grad=True)
    val, gp = T.slice_per_corner(vol, p, with_grad=True)
    np.testing.assert_allclose(val, ref, atol=1e-12)
    # away from cell faces the one-sided derivative is the derivative
    frac = p - np.floor(p)
    ok = np.all((frac > 1e-6) & (frac < 1 - 1e-6), axis=1)
    np.testing.assert_allclose((gp * w[:, None])[ok], gref[ok], atol=1e-10)


def test_planar_operator_and_uv_gradient_match_torch_autograd():
    rng = np.random.default_rng(2)
    vol_ri = rng.standard_normal((2, 2, 5, 6, 7))
    N, K = 3, 50
    u = rng.standard_normal((N, 3)) * 2.0
    v = rng.standard_normal((N, 3)) * 2.0
    fx, fy = rng.uniform(-1, 1, K), rng.uniform(-1, 1, K)
    cen = (3.0, 2.5, 2.0)
    out = T.fourier_slice_3d(vol_ri, u, v, fx, fy, cen)
    go = rng.standard_normal(out.shape)
    gu, gv = T.fourier_slice_3d_grad_uv(vol_ri, u, v, fx, fy, cen, go)
    # torch autograd of grid_sample with the grid built from u, v
    ut = torch.tensor(u, requires_grad=True)
    vt = torch.tensor(v, requires_grad=True)
    B, _, D, H, W = vol_ri.shape
    f64 = torch.float64
    p = torch.tensor(cen, dtype=f64)[None, None] + torch.tensor(fx)[None, :, None] * ut[:, None] \
        + torch.tensor(fy)[None, :, None] * vt[:, None]
    g = p * torch.tensor([2.0 / (W - 1), 2.0 / (H - 1), 2.0 / (D - 1)], dtype=f64) - 1.0
    s = torch.nn.functional.grid_sample(torch.tensor(vol_ri), g.view(1, N, K, 1, 3).expand(B, -1, -1, -1, -1),
                                        mode="bilinear", padding_mode="zeros", align_corners=True)
    s = s.squeeze(-1).permute(0, 2, 1, 3)  # (B, N, 2, K)
    np.testing.assert_allclose(out, s.detach().numpy(), atol=1e-12)
    (s * torch.tensor(go)).sum().backward()
    np.testing.assert_allclose(gu, ut.grad.numpy(), atol=1e-9)
    np.testing.assert_allclose(gv, vt.grad.numpy(), atol=1e-9)


def test_hp_filter_closed_form():
    hp = T.hp_filter(9, 9, 2.0)  # odd size: the grid hits 0 at the centre
    assert hp[4, 4] == 0.0
    assert abs(hp[0, 0] - (1.0 - math.exp(-(4.5 ** 2 * 2) / 8.0))) < 1e-15
    assert T.hp_filter(64, 64, 4.0)[0, 0] == 1.0  # exp(-256) underflows against 1


def test_self_target_gives_zero_loss_and_fd_gradient():
    v = _vol((8, 9, 10))
    S = T.t_spectrum(v)
    R = 15
    pose = np.array([2.0, -1.5, 1.0, 0.7, -0.4, 0.3])
    tgt = T.extract_slice_spectrum(S, pose, "ap", R) * T.hp_filter(R, R, 4.0)
    E = np.sum(np.abs(tgt) ** 2)
    # NCC of a slice with itself is E / (E + 1e-8): the loss is exactly the eps share
    assert abs(T.spectral_ncc_core(S, pose[None], tgt, "ap") - 1e-8 / (E + 1e-8)) < 1e-14
    q = np.array([[4.0, 2.5, -3.0, 1.2, -0.8, 0.0], [-3.0, 1.0, 2.0, -1.0, 0.5, 0.4]])
    for view in ("ap", "lateral"):
        L, g, _ = T.spectral_ncc_core_and_grad(S, q, tgt, view)
        assert 0.0 <= L <= 2.0
        h = 1e-3
        for b in range(2):
            for k in range(6):
                qp, qm = q.copy(), q.copy()
                qp[b, k] += h
                qm[b, k] -= h
                fd = (T.spectral_ncc_core(S, qp, tgt, view) - T.spectral_ncc_core(S, qm, tgt, view)) / (2 * h)
                assert abs(fd - g[b, k]) <= 2e-6 * max(1.0, abs(g).max()), (view, b, k, fd, g[b, k])
        assert np.all(g[:, 5] == 0.0)  # tz does not enter the T path


# ---------------------------------------------------------------------------
# host side of the fused T path: item geometry and the moment combine
# ---------------------------------------------------------------------------
class _FakePacked:
    """Minimal stand-in for PackedSpectrum's descriptor table (no device)."""

    def __init__(self, shape):
        from fvr import _lib
        Z, Y, X = shape
        self.index = {"zyx": 0, "zxy": 1}
        self.descs = np.zeros(2, dtype=_lib.LAYOUT_DTYPE)
        for k, n in enumerate(([Z, Y, X], [Z, X, Y])):
            d = self.descs[k]
            d["n"] = n
            d["sdim"] = [n[0] + 3, n[1] + 2, n[2] + 2]
            d["elem_bytes"] = 8
            d["row_bytes"] = 8 * (n[2] + 2)
            d["plane_bytes"] = d["row_bytes"] * (n[1] + 2)

    def desc(self, k):
        return self.descs[k]


def _moments_by_definition(S3, pose, view, tgt, R, sigma, perm_order):
    """The 16 sums the fused kernel accumulates, from the oracle's interpolation (float64)."""
    p, f, u, v, du, dv = T._slice_points(pose, view, S3.shape, R)
    Ir, gr = T.slice_per_corner(S3.real, p, with_grad=True)
    Ii, gi = T.slice_per_corner(S3.imag, p, with_grad=True)
    I = Ir + 1j * Ii
    gI = (gr + 1j * gi)[..., perm_order]  # layout axes
    span = R * 1.0 + 1e-6
    ph = np.exp(-2j * np.pi * (f[:, None] * pose[3] + f[None, :] * pose[4]) / span)
    hw = T.hp_filter(R, R, sigma) if sigma else np.ones((R, R))
    Wt = ph * np.conj(tgt) * hw
    J = I * hw ** 2
    cw = np.broadcast_to(f[None, :], I.shape)
    rw = np.broadcast_to(f[:, None], I.shape)
    IW = I * Wt
    m = [np.sum(IW.real), np.sum(np.real(I * np.conj(J))), np.sum(cw * IW.imag), np.sum(rw * IW.imag)]
    for wgt in (cw, rw):
        m += [np.sum(wgt * np.real(gI[..., a] * Wt)) for a in range(3)]
    for wgt in (cw, rw):
        m += [np.sum(wgt * np.real(gI[..., a] * np.conj(J))) for a in range(3)]
    return np.array(m)


@pytest.mark.parametrize("view", ["ap", "lateral"])
def test_t_items_geometry_and_combine_reproduce_oracle(view):
    from fvr import _items, _lib
    from fvr.torch_engine import combine_t
    v = _vol((8, 9, 10))
    S3 = T.t_spectrum(v)
    R, sigma = 13, 4.0
    packed = _FakePacked(S3.shape)
    poses = np.array([[3.0, -2.0, 1.0, 0.8, -0.6, 0.2], [-7.0, 5.0, 4.0, -1.1, 0.9, 0.0]])
    tgt = T.extract_slice_spectrum(S3, np.array([1.0, 0.5, -0.5, 0.2, 0.1, 0.0]), "ap", R) * T.hp_filter(R, R, sigma)
    tmpl = _items.t_template(packed, view, R, 0x1000, _lib.ITEM_ADJOINT, 64, sigma)
    items, jac, n_tiles = _items.t_items(tmpl, S3.shape, poses, view, R, 1.0, 64)
    assert n_tiles == len(poses) * 1
    assert np.all(items["flags"] == (_lib.ITEM_ADJOINT | _lib.ITEM_HP))
    lay = 1 if view == "ap" else 0
    perm = [0, 2, 1] if lay == 1 else [0, 1, 2]
    order = [[2, 1, 0][a] for a in perm]
    for b, pose in enumerate(poses):
        # fixed-point positions (lo_int = 0: +1 shift) reproduce centre + fx_i u + fy_j v
        p, *_ = T._slice_points(pose, view, S3.shape, R)
        ii, jj = np.meshgrid(np.arange(R), np.arange(R), indexing="ij")
        it = items[b]
        fxp = (it["c"][None, None, :].astype(np.float64) + jj[..., None] * it["ustep"].astype(np.float64)
               + ii[..., None] * it["vstep"].astype(np.float64) - _items.SNAP_FX) / 2.0 ** 32 - 1.0
        np.testing.assert_allclose(fxp, p[..., order], atol=1e-8)
        mom = _moments_by_definition(S3, pose, view, tgt, R, sigma, order)
        e_t = np.array([np.sum(np.abs(tgt) ** 2)])
        ncc, d = combine_t(mom[None], jac[b:b + 1], e_t, _items.t_phase_span(R, 1.0))
        L1, g1, ncc1 = T.spectral_ncc_core_and_grad(S3, pose[None], tgt, view, True, sigma)
        assert abs(ncc[0] - ncc1[0]) < 1e-12
        np.testing.assert_allclose(-d[0], g1[0], atol=1e-12 * max(1.0, np.abs(g1).max()))


@pytest.mark.parametrize("shape", [(12, 12), (10, 14)])
def test_image_spectral_ncc_loss_parseval_invariants_and_fd(shape):
    """Oracle spectral_ncc_loss (torch_engine.py:309-337) pinned without retyping it:
    Parseval in the image domain (high-pass off: NCC = sum p t / (|p| |t|); high-pass on:
    the same for the images filtered by an inverse FFT of hp * F), NCC(p, p) = 1,
    NCC(p, -p) = -1, scale invariance, the batch mean, and the gradient w.r.t. the
    predicted image against FP64 central differences."""
    H, W = shape
    rng = np.random.default_rng(7)
    p = rng.standard_normal((3, 1, H, W))
    t = rng.standard_normal((3, 1, H, W))
    hw = H * W
    # high-pass off: Parseval, with the reference's 1e-8 added to sqrt(Ep Et) = H W |p| |t|
    ncc = np.sum(p * t, axis=(-2, -1)) / (np.sqrt(np.sum(p * p, axis=(-2, -1)) * np.sum(t * t, axis=(-2, -1)))
                                         + 1e-8 / hw)
    assert abs(T.spectral_ncc_loss(p, t, high_pass=False) - (1.0 - ncc.mean())) < 1e-12
    # high-pass on: images filtered in the Fourier domain, inner products in the image domain

    def filt(x):
        F = np.fft.fftshift(np.fft.fft2(x), axes=(-2, -1)) * T.hp_filter(H, W, 4.0)
        return np.fft.ifft2(np.fft.ifftshift(F, axes=(-2, -1)))

    pf, tf = filt(p), filt(t)
    ncc_hp = np.sum(np.real(pf * np.conj(tf)), axis=(-2, -1)) / (
        np.sqrt(np.sum(np.abs(pf) ** 2, axis=(-2, -1)) * np.sum(np.abs(tf) ** 2, axis=(-2, -1))) + 1e-8 / hw)
    assert abs(T.spectral_ncc_loss(p, t) - (1.0 - ncc_hp.mean())) < 1e-12
    # invariants (up to the 1e-8 in the denominator: relative 1e-8 / (H W |p|^2) here)
    assert abs(T.spectral_ncc_loss(p, p)) < 1e-10
    assert abs(T.spectral_ncc_loss(p, -p) - 2.0) < 1e-10
    assert abs(T.spectral_ncc_loss(3.5 * p, 0.25 * t) - T.spectral_ncc_loss(p, t)) < 1e-12
    singles = [T.spectral_ncc_loss(p[b:b + 1], t[b:b + 1]) for b in range(3)]
    assert abs(T.spectral_ncc_loss(p, t) - np.mean(singles)) < 1e-12
    # gradient w.r.t. the predicted image
    for hpass in (True, False):
        L, g = T.spectral_ncc_loss(p, t, high_pass=hpass, with_grad=True)
        h = 1e-6
        fd = np.zeros_like(p)
        for idx in np.ndindex(p.shape):
            pp, pm = p.copy(), p.copy()
            pp[idx] += h
            pm[idx] -= h
            fd[idx] = (T.spectral_ncc_loss(pp, t, high_pass=hpass) - T.spectral_ncc_loss(pm, t, high_pass=hpass)) / (2 * h)
        assert np.abs(g - fd).max() <= 1e-7 * np.abs(fd).max(), np.abs(g - fd).max()
