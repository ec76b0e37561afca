"""Pins the convention-N oracle (oracle/diff_fvr_np.py) on the CPU.

Each check is independent of the formula it pins: the reference's own known
answers, library routines called directly, closed forms, brute force and
finite differences.
"""

import math

import numpy as np
import torch
import pytest
from scipy import fft
from scipy.ndimage import map_coordinates

from oracle import diff_fvr_np as O


def _phantom24():
    # the reference's test fixture (tests/test_physical_units.py:21-24)
    z, y, x = np.indices((24,) * 3)
    return (np.exp(-((z - 10) ** 2 + (y - 11) ** 2 + (x - 10) ** 2) / 8)
            + .7 * np.exp(-((z - 15) ** 2 + (y - 15) ** 2 + (x - 15) ** 2) / 4))


@pytest.fixture(scope="module")
def cube():
    vol = np.zeros((20, 20, 20))
    vol[6:14, 5:13, 7:15] = 1.0
    vol[3:6, 9:12, 4:8] = 2.0
    return vol, O.make_spectrum(vol)


def test_known_answer_mm_shift_is_np_roll():
    # tests/test_physical_units.py:28-34: 4/-2 mm at 2 mm spacing == roll by (-1, 2)
    vol = _phantom24()
    f3 = O.make_spectrum(vol)
    for view in ("ap", "lateral"):
        img = O.render_drr(f3, vol.shape, view=view, crop_to_roi=False)
        shifted = O.render_drr(f3, vol.shape, tx=4 / 2.0, ty=-2 / 2.0, view=view, crop_to_roi=False)
        np.testing.assert_allclose(shifted, np.roll(img, (-1, 2), axis=(0, 1)), atol=1e-9)


def test_sinc2_window_closed_form():
    n, p = 10, 20
    i = np.arange(n)
    x = (i - n / 2) / p
    expect = np.array([1.0 / max((math.sin(math.pi * v) / (math.pi * v)) ** 2 if v else 1.0, 0.05) for v in x])
    np.testing.assert_allclose(O.sinc2_correction(n, p), expect, rtol=1e-14)


def test_spectrum_of_delta_is_pure_phase_ramp():
    vol = np.zeros((8, 8, 8))
    vol[3, 5, 2] = 1.0
    f3 = O.make_spectrum(vol, premultiply=False)
    pz = 16
    k = np.arange(pz) - pz // 2
    # padded index of the voxel relative to the centre pz/2 after the (p-n)//2 offset
    dz, dy, dx = 3 + 4 - 8, 5 + 4 - 8, 2 + 4 - 8
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    expect = np.exp(-2j * np.pi * (kz * dz + ky * dy + kx * dx) / pz)
    np.testing.assert_allclose(f3, expect, atol=1e-12)


def test_fourier_slice_theorem_at_identity(cube):
    # the identity AP slice samples lattice points exactly: it equals the 2D
    # FFT of the y-projection of the premultiplied padded volume
    vol, f3 = cube
    s = O.extract_frequency_slice(f3, view="ap")
    proj = O.make_padded_volume(vol).sum(axis=1)
    np.testing.assert_allclose(s, fft.fftshift(fft.fft2(fft.ifftshift(proj))), atol=1e-9)
    s = O.extract_frequency_slice(f3, view="lateral")
    proj = O.make_padded_volume(vol).sum(axis=2)
    np.testing.assert_allclose(s, fft.fftshift(fft.fft2(fft.ifftshift(proj))), atol=1e-9)


def test_rotation_orthonormal_and_90_degree_cases():
    rng = np.random.default_rng(0)
    for _ in range(5):
        R = O.euler_to_rotation(*rng.uniform(-180, 180, 3))
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-12
    # yaw 90: rotates within (Y, X) -> x-axis maps to -y
    R = O.euler_to_rotation(90, 0, 0)
    np.testing.assert_allclose(R @ [0, 0, 1], [0, -1, 0], atol=1e-15)
    u, v = O.euler_to_basis(0, 0, 0, "ap")
    np.testing.assert_allclose(u, [0, 0, 1])
    np.testing.assert_allclose(v, [1, 0, 0])


def test_trilinear_value_matches_scipy_including_edges():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((5, 6, 7)) + 1j * rng.standard_normal((5, 6, 7))
    c = np.stack([rng.uniform(-1.5, s + 0.5, 4000) for s in a.shape])
    c[:, :400] = np.round(c[:, :400])          # exact lattice / boundary coordinates
    c[0, 400:500] = 4.0                        # exactly n-1 on one axis
    c[2, 500:600] = -1e-12                     # just below 0
    I, _ = O.trilinear_value_and_grad(a, c)
    ref = (map_coordinates(a.real, c, order=1, mode="constant", cval=0.0)
           + 1j * map_coordinates(a.imag, c, order=1, mode="constant", cval=0.0))
    np.testing.assert_allclose(I, ref, atol=1e-13)


def test_trilinear_gradient_brute_force_and_fd():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((6, 6, 6)) + 1j * rng.standard_normal((6, 6, 6))
    c = np.stack([rng.uniform(0.05, 4.95, 300) for _ in range(3)])
    _, dI = O.trilinear_value_and_grad(a, c)
    h = 1e-6
    for ax in range(3):
        cp, cm = c.copy(), c.copy()
        cp[ax] += h
        cm[ax] -= h
        fd = (O.trilinear_value_and_grad(a, cp)[0] - O.trilinear_value_and_grad(a, cm)[0]) / (2 * h)
        # away from cell faces the interpolant is smooth
        t = c[ax] - np.floor(c[ax])
        ok = (t > 1e-3) & (t < 1 - 1e-3)
        np.testing.assert_allclose(dI[ax][ok], fd[ok], atol=1e-7)
    # brute-force per-corner derivative at one point
    p = np.array([2.3, 1.7, 3.9])
    i0 = np.floor(p).astype(int)
    t = p - i0
    gx = 0
    for dz in (0, 1):
        for dy in (0, 1):
            wz = t[0] if dz else 1 - t[0]
            wy = t[1] if dy else 1 - t[1]
            gx += wz * wy * (a[i0[0] + dz, i0[1] + dy, i0[2] + 1] - a[i0[0] + dz, i0[1] + dy, i0[2]])
    assert abs(O.trilinear_value_and_grad(a, p[:, None])[1][2, 0] - gx) < 1e-13


def test_ncc_properties(cube):
    _, f3 = cube
    s = O.extract_frequency_slice(f3, 2.0, -1.0, 0.5, view="ap")
    assert abs(O.compute_spectral_ncc(s, s) - 1.0) < 1e-12
    assert abs(O.compute_spectral_ncc(3.5 * s, 0.25 * s) - 1.0) < 1e-12
    assert O.compute_spectral_ncc(s, np.zeros_like(s)) == 0.0
    # Parseval: spectral NCC of two spectra == spatial correlation of their images
    t = O.extract_frequency_slice(f3, -1.0, 1.0, 0.0, tx=0.7, view="ap")
    a = fft.ifft2(fft.ifftshift(s))
    b = fft.ifft2(fft.ifftshift(t))
    spatial = np.real(np.vdot(b, a)) / (np.linalg.norm(a) * np.linalg.norm(b))
    assert abs(O.compute_spectral_ncc(s, t) - spatial) < 1e-12


@pytest.fixture(scope="module")
def bip(cube):
    _, f3 = cube
    tp = np.array([1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
    T_ap = O.extract_frequency_slice(f3, *tp[:3], tx=tp[3], ty=tp[5], view="ap")
    T_la = O.extract_frequency_slice(f3, *tp[:3], tx=tp[4], ty=tp[5], view="lateral")
    return f3, T_ap, T_la


@pytest.mark.parametrize("pose", [[3.3, 1.1, -2.2, 2.1, 1.5, -1.0], [-6.0, 4.0, 2.0, -3.0, 0.5, 2.5]])
def test_analytic_gradient_matches_central_differences(bip, pose):
    f3, T_ap, T_la = bip
    pose = np.array(pose)
    _, g = O.biplanar_loss_and_grad_analytic(f3, pose, T_ap, T_la)
    for h in (1e-3, 1e-5):
        fd = np.zeros(6)
        for i in range(6):
            pp, pm = pose.copy(), pose.copy()
            pp[i] += h
            pm[i] -= h
            fd[i] = (O.biplanar_loss(f3, pp, T_ap, T_la) - O.biplanar_loss(f3, pm, T_ap, T_la)) / (2 * h)
        assert np.max(np.abs(fd - g)) <= 1e-3 * np.max(np.abs(g)), (h, fd, g)


def test_reference_scheme_translation_terms_equal_analytic(bip):
    f3, T_ap, T_la = bip
    pose = np.array([2.0, -1.0, 1.0, 1.5, -0.5, 0.8])
    L1, g_fd = O.biplanar_loss_and_gradients(f3, pose, T_ap, T_la)
    L2, g_an = O.biplanar_loss_and_grad_analytic(f3, pose, T_ap, T_la)
    assert abs(L1 - L2) < 1e-14
    np.testing.assert_allclose(g_fd[3:], g_an[3:], rtol=1e-9, atol=1e-14)
    # the reference's 0.2-degree central differences approximate the exact rotations
    assert np.max(np.abs(g_fd[:3] - g_an[:3])) < 0.5 * np.max(np.abs(g_an[:3])) + 1e-6
    # the FD scheme itself: shrinking delta converges to the exact gradient
    _, g_small = O.biplanar_loss_and_gradients(f3, pose, T_ap, T_la, delta_rot_deg=1e-4)
    np.testing.assert_allclose(g_small[:3], g_an[:3], rtol=1e-4, atol=1e-9)


def test_mtre_reference_closed_forms():
    # tests/test_physical_units.py:52-59
    o = (0.0,) * 6
    assert abs(O.compute_mtre_6dof(o, (0, 0, 0, 3, 0, 0), (16,) * 3, spacing=(2.0,) * 3) - 6.0) < 1e-12
    assert abs(O.compute_mtre_6dof(o, (0, 0, 0, 3, 0, 0), (16,) * 3, spacing=(2.0,) * 3,
                                   translation_units="mm") - 3.0) < 1e-12
    assert abs(O.compute_mtre_6dof(o, (90, 0, 0, 0, 0, 0), (16,) * 3, spacing=(2.0,) * 3) - 16.0) < 1e-12


def test_adam_closed_form_on_quadratic():
    target = np.array([1.0, -2.0, 0.5, 3.0, -1.0, 2.0])

    def f(p):
        return float(np.sum((p - target) ** 2)), 2 * (p - target)

    res = O.adam_register(f, np.zeros(6), max_iter=1, tol=0.0)
    # first Adam step moves every coordinate by exactly lr * sign(grad) (m_hat/sqrt(v_hat) = +-1)
    step1 = res["trajectory"][0][2]
    assert np.allclose(step1, 0.0)
    res2 = O.adam_register(f, np.zeros(6), max_iter=2, tol=0.0)
    p1 = res2["trajectory"][1][2]
    lr = np.array([0.25] * 3 + [0.35] * 3)
    np.testing.assert_allclose(p1, lr * np.sign(target), rtol=1e-6)


def test_adam_multistart_options_equal_torch_adam_with_lambdalr():
    """exp30's optimiser: torch.optim.Adam(betas=(0.8, 0.99)) per parameter group (rotations,
    translations), gradient clamp, LambdaLR cosine factor -- in float64 on a fixed sequence."""
    rng = np.random.default_rng(5)
    G = rng.standard_normal((30, 6)) * np.array([0.4] * 3 + [4.0] * 3)
    p0 = rng.standard_normal(6)
    k = {"n": 0}

    def f(p):
        g = G[k["n"]]
        k["n"] += 1
        return 1.0, g

    T = 20
    res = O.adam_register(f, p0, lr_rot=0.02, lr_trans=2.0, max_iter=30, tol=0.0, beta1=0.8, beta2=0.99,
                          limits=None, gclip=(0.25, 2.5), cosine_t=T)
    rot = torch.tensor(p0[:3], dtype=torch.float64, requires_grad=True)
    tr = torch.tensor(p0[3:], dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([{"params": [rot], "lr": 0.02}, {"params": [tr], "lr": 2.0}], betas=(0.8, 0.99))
    sched = torch.optim.lr_scheduler.LambdaLR(opt, lambda s: 0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * min(s, T) / T)))
    for i in range(30):
        np.testing.assert_allclose(np.concatenate([rot.detach().numpy(), tr.detach().numpy()]),
                                   res["trajectory"][i][2], rtol=1e-12, atol=1e-12)
        rot.grad = torch.tensor(G[i, :3]).clamp(-0.25, 0.25)
        tr.grad = torch.tensor(G[i, 3:]).clamp(-2.5, 2.5)
        opt.step()
        sched.step()


def test_prepare_spectrum_identity_roundtrip(cube):
    vol, f3 = cube
    drr = O.render_drr(f3, vol.shape, view="ap", crop_to_roi=False)
    # scaled inverse of render_drr for a non-negative image (no ReLU clipping)
    spec = O.prepare_spectrum(drr, *drr.shape)
    s = O.extract_frequency_slice(f3, view="ap")
    scale = np.prod(f3.shape) / np.prod(drr.shape)
    # the identity projection of a non-negative box phantom has no negative lobe, so ReLU is inactive
    assert np.min(np.real(fft.fftshift(fft.ifft2(fft.ifftshift(s))))) >= -1e-9
    np.testing.assert_allclose(spec, s * scale, atol=1e-8 * np.abs(s).max() * scale)
