"""GPU parity of the convention-T path (TorchDiffFVREngine, fourier_slice_3d) against the
FP64 oracle (oracle/torch_engine_np.py), plus the reference's own torch-engine checks
(tests/test_torch_engine.py of the reference) on the HIP path.

Tolerances: the generic operator computes sample positions in float32 like the
reference's kernel (triton_slice.py:83-85), so its outputs are compared at
1e-5 of max|ref| (f32 storage) and its u/v gradients at 1e-4; the fused path
uses 32.32 fixed-point positions from float64 geometry: loss within 1e-5
relative, gradient within 1e-4 of max|grad| (the target bound is 1e-3 relative).
"""

import numpy as np
import pytest
import torch

from oracle import torch_engine_np as T

pytestmark = pytest.mark.gpu


def _phantom(n=16, seed=7):
    from fvr.phantom import ellipsoid_phantom
    return ellipsoid_phantom(n, n_ellipsoids=6, seed=seed).astype(np.float32)


@pytest.fixture(scope="module")
def eng():
    from fvr import TorchDiffFVREngine
    vol = _phantom()
    e = TorchDiffFVREngine(torch.from_numpy(vol), spacing=(1.0, 1.0, 1.0), pad_factor=2.0)
    S3 = T.t_spectrum(vol.astype(np.float64))
    return e, S3


def test_spectrum_and_planar_buffer(eng):
    e, S3 = eng
    assert tuple(e.pad_shape) == S3.shape
    f = e.f_vol_5d[0].double().cpu().numpy()
    np.testing.assert_allclose(f[0] + 1j * f[1], S3, atol=2e-6 * np.abs(S3).max())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.bfloat16])
def test_fourier_slice_3d_forward_and_uv_gradient(dtype):
    from fvr import fourier_slice_3d
    rng = np.random.default_rng(11)
    vol = rng.standard_normal((2, 2, 6, 7, 8)).astype(np.float32)
    N, K = 3, 333
    u = (rng.standard_normal((N, 3)) * 3).astype(np.float32)
    v = (rng.standard_normal((N, 3)) * 3).astype(np.float32)
    fx = rng.uniform(-1, 1, K).astype(np.float32)
    fy = rng.uniform(-1, 1, K).astype(np.float32)
    fx[:3], fy[:3] = [-1, 0, 1], [1, 0, -1]
    cen = (3.5, 3.0, 2.5)
    dev = torch.device("cuda", 0)
    vt = torch.from_numpy(vol).to(dev, dtype)
    vq = vt.float().cpu().numpy().astype(np.float64)  # the stored (rounded) values
    ut = torch.from_numpy(u).to(dev).requires_grad_(True)
    vv = torch.from_numpy(v).to(dev).requires_grad_(True)
    out = fourier_slice_3d(vt, ut, vv, torch.from_numpy(fx).to(dev), torch.from_numpy(fy).to(dev), cen)
    assert out.shape == (2, N, 2, K) and out.dtype == dtype
    ref = T.fourier_slice_3d(vq, u.astype(np.float64), v.astype(np.float64), fx, fy, cen)
    tol = {torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1.6e-2}[dtype]
    assert np.max(np.abs(out.detach().float().cpu().numpy() - ref)) <= tol * np.abs(ref).max()
    go = rng.standard_normal(out.shape).astype(np.float32)
    (out.float() * torch.from_numpy(go).to(dev)).sum().backward()
    gu, gv = T.fourier_slice_3d_grad_uv(vq, u.astype(np.float64), v.astype(np.float64), fx, fy, cen,
                                        torch.from_numpy(go).to(dtype).float().numpy())
    gtol = {torch.float32: 1e-4, torch.float16: 2e-3, torch.bfloat16: 1e-2}[dtype]
    for got, want in ((ut.grad, gu), (vv.grad, gv)):
        assert np.max(np.abs(got.cpu().numpy() - want)) <= gtol * np.abs(want).max()


def test_fourier_slice_3d_refuses_volume_gradients_and_cpu():
    from fvr import fourier_slice_3d
    dev = torch.device("cuda", 0)
    vol = torch.zeros((1, 2, 4, 4, 4), device=dev, requires_grad=True)
    u = torch.zeros((1, 3), device=dev)
    f = torch.zeros(4, device=dev)
    with pytest.raises(NotImplementedError):
        fourier_slice_3d(vol, u, u, f, f, (1.5, 1.5, 1.5))
    with pytest.raises(RuntimeError):
        fourier_slice_3d(torch.zeros((1, 2, 4, 4, 4)), u.cpu(), u.cpu(), f.cpu(), f.cpu(), (1.5, 1.5, 1.5))


@pytest.mark.parametrize("view", ["ap", "lateral"])
def test_forward_drr_matches_oracle(eng, view):
    e, S3 = eng
    poses = np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [3.0, -2.0, 1.0, 1.5, -1.0, 0.0]])
    out = e(torch.from_numpy(poses).float(), view=view, out_res=24)
    assert out.shape == (2, 1, 24, 24)
    for b in range(2):
        ref = T.forward_drr(S3, poses[b], view, 24)
        got = out[b, 0].double().cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-4 * np.abs(ref).max()


def _target(e, S3, pose, view, R, high_pass=True):
    img = T.forward_drr(S3, np.asarray(pose, np.float64), view, R)
    ft = e.prepare_spectral_target(torch.from_numpy(img).float()[None, None], high_pass=high_pass)
    return ft, T.prepare_spectral_target(img, high_pass)


@pytest.mark.parametrize("view", ["ap", "lateral"])
@pytest.mark.parametrize("high_pass", [True, False])
def test_fused_spectral_ncc_loss_and_gradient(eng, view, high_pass):
    e, S3 = eng
    R = 24
    ft, ft_np = _target(e, S3, [1.0, -0.5, 0.5, 0.4, -0.3, 0.0], view, R, high_pass)
    # prepare_spectral_target (torch_engine.py:255-267: the image cast to float32, FFT on the
    # device) against the oracle's float64 preparation of the same image
    assert tuple(ft.shape) == (1, 1, R, R)
    assert np.abs(ft[0, 0].cpu().numpy() - ft_np).max() <= 1e-5 * np.abs(ft_np).max()
    poses = np.array([[3.0, -2.0, 1.0, 1.5, -1.0, 0.3], [-4.0, 2.5, -1.5, -0.8, 1.2, 0.0],
                      [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]])
    P = torch.from_numpy(poses).float().requires_grad_(True)
    loss = e.forward_spectral_ncc(P, ft, view=view, high_pass=high_pass)
    loss.backward()
    # the oracle uses the float32 target the engine holds
    t32 = ft[0, 0].to(torch.complex128).cpu().numpy()
    L, g, _ = T.spectral_ncc_core_and_grad(S3, poses.astype(np.float32).astype(np.float64), t32, view, high_pass)
    assert abs(float(loss) - L) <= 1e-5 * max(abs(L), 1e-3)
    gg = P.grad.double().numpy()
    assert np.max(np.abs(gg - g)) <= 1e-4 * np.abs(g).max() + 1e-9, (gg, g)
    assert np.all(gg[:, 5] == 0.0)


def test_fused_gradient_matches_torch_autograd_of_generic_path(eng):
    """The fused adjoint against torch autograd through the generic slice operator and the
    reference's torch ops (torch_engine.py:269-290 structure)."""
    e, S3 = eng
    R, view = 24, "ap"
    ft, _ = _target(e, S3, [1.0, -0.5, 0.5, 0.4, -0.3, 0.0], view, R)
    poses = torch.tensor([[3.0, -2.0, 1.0, 1.5, -1.0, 0.3], [-4.0, 2.5, -1.5, -0.8, 1.2, 0.0]])
    P1 = poses.clone().requires_grad_(True)
    e.forward_spectral_ncc(P1, ft, view=view).backward()
    P2 = poses.clone().cuda().requires_grad_(True)
    f_pred = e._extract_slice_spectrum(P2, view=view, out_res=R).unsqueeze(1)
    f_pred = f_pred * e._get_hp_filter(R, R, 4.0, P2.device, torch.float32)
    cross = torch.sum(torch.real(f_pred * torch.conj(ft)), dim=(-2, -1))
    ep = torch.sum(f_pred.real.square() + f_pred.imag.square(), dim=(-2, -1))
    et = torch.sum(ft.real.square() + ft.imag.square(), dim=(-2, -1))
    loss2 = 1.0 - torch.mean(cross / (torch.sqrt(ep * et) + 1e-8))
    loss2.backward()
    g1, g2 = P1.grad.numpy(), P2.grad.cpu().numpy()
    assert np.max(np.abs(g1 - g2)) <= 2e-3 * np.abs(g2).max(), (g1, g2)


@pytest.mark.parametrize("dtype,tol_l,tol_g", [(torch.float32, 1e-5, 1e-4), (torch.float64, 1e-10, 1e-9)])
@pytest.mark.parametrize("high_pass", [True, False])
def test_image_spectral_ncc_loss_matches_oracle(eng, dtype, tol_l, tol_g, high_pass):
    """spectral_ncc_loss (torch_engine.py:309-337) on device DRRs of three poses against
    the oracle's FP64 loss and its exact gradient w.r.t. the predicted images (autograd
    through rocFFT on the device): float32 inputs within 1e-5 (loss) and 1e-4 of
    max|grad|, float64 inputs within 1e-10 and 1e-9."""
    e, _ = eng
    poses = torch.tensor([[3.0, -2.0, 1.0, 1.5, -1.0, 0.3], [-4.0, 2.5, -1.5, -0.8, 1.2, 0.0],
                          [0.5, 0.5, -0.5, 0.2, 0.1, 0.0]])
    with torch.no_grad():
        pred = e(poses, view="ap", out_res=32).to(dtype)
        tgt = e(torch.tensor([[1.0, -0.5, 0.5, 0.4, -0.3, 0.0]]), view="ap", out_res=32).to(dtype)
    x = pred.clone().requires_grad_(True)
    loss = e.spectral_ncc_loss(x, tgt, high_pass=high_pass)
    loss.backward()
    L, g = T.spectral_ncc_loss(pred.double().cpu().numpy(), tgt.double().cpu().numpy(), high_pass=high_pass,
                               with_grad=True)
    assert abs(float(loss) - L) <= tol_l, (float(loss), L)
    gg = x.grad.double().cpu().numpy()
    assert np.abs(gg - g).max() <= tol_g * np.abs(g).max(), np.abs(gg - g).max()


def test_per_pose_targets_and_batched_equals_single(eng):
    e, S3 = eng
    R = 24
    fa, _ = _target(e, S3, [1.0, 0.0, 0.0, 0.0, 0.0, 0.0], "ap", R)
    fb, _ = _target(e, S3, [-1.0, 0.5, 0.0, 0.5, 0.0, 0.0], "ap", R)
    poses = torch.tensor([[2.0, -1.0, 0.5, 0.3, 0.2, 0.0], [-2.0, 1.0, 0.0, -0.4, 0.1, 0.0]])
    both = torch.cat([fa, fb], 0)
    Lb = float(e.forward_spectral_ncc(poses, both, view="ap"))
    L0 = float(e.forward_spectral_ncc(poses[:1], fa, view="ap"))
    L1 = float(e.forward_spectral_ncc(poses[1:], fb, view="ap"))
    assert abs(Lb - 0.5 * (L0 + L1)) < 1e-6


def test_non_finite_pose_gives_nan(eng):
    """forward_spectral_ncc with a NaN angle or in-plane translation in one pose: the loss
    (1 - mean NCC) is NaN and that pose's gradient is NaN; the other poses' gradients are the
    ones of the same batch with that pose finite, bit for bit (tz does not enter
    convention T, so a NaN there changes nothing)."""
    e, S3 = eng
    R = 24
    ft, _ = _target(e, S3, [1.0, -0.5, 0.5, 0.4, -0.3, 0.0], "ap", R)
    base = torch.tensor([[3.0, -2.0, 1.0, 1.5, -1.0, 0.3], [-4.0, 2.5, -1.5, -0.8, 1.2, 0.0],
                         [0.5, 0.5, -0.5, 0.2, 0.1, 0.0]])

    def run(p):
        x = p.clone().requires_grad_(True)
        loss = e.forward_spectral_ncc(x, ft, view="ap")
        loss.backward()
        return float(loss), x.grad.double().numpy()

    L0, g0 = run(base)
    for col in (0, 3):
        bad = base.clone()
        bad[1, col] = float("nan")
        L1, g1 = run(bad)
        assert np.isnan(L1) and np.all(np.isnan(g1[1]))
        np.testing.assert_array_equal(g1[[0, 2]], g0[[0, 2]])
    tz = base.clone()
    tz[1, 5] = float("nan")
    L2, g2 = run(tz)
    assert L2 == L0 and np.array_equal(g2, g0)


# --- the reference's engine checks (tests/test_torch_engine.py) on the HIP path --------------
def test_reference_forward_batch_and_autograd(eng):
    e, _ = eng
    poses = torch.tensor([[0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [5.0, 0.0, 0.0, 0.0, 0.0, 0.0],
                          [0.0, 5.0, 0.0, 0.0, 0.0, 0.0], [2.0, -3.0, 1.5, 4.0, -2.0, 1.0]])
    for view in ("ap", "lateral"):
        d = e(poses, view=view, out_res=32)
        assert d.shape == (4, 1, 32, 32) and bool(torch.all(torch.isfinite(d))) and bool(torch.all(d >= 0))
    pose = torch.tensor([[3.0, -2.0, 1.0, 5.0, -3.0, 2.0]], requires_grad=True)
    torch.sum(e(pose, view="ap", out_res=32) ** 2).backward()
    assert pose.grad.shape == (1, 6) and bool(torch.all(torch.isfinite(pose.grad)))
    assert not bool(torch.all(pose.grad == 0))


def test_reference_adam_descends_with_fused_loss(eng):
    e, _ = eng
    with torch.no_grad():
        tgt = e(torch.zeros((1, 6)), view="ap", out_res=32)
        ft = e.prepare_spectral_target(tgt, high_pass=True)
    est = torch.tensor([[2.5, -1.5, 0.5, 3.0, -2.5, 0.0]], requires_grad=True)
    opt = torch.optim.Adam([est], lr=0.2)
    losses = []
    for _ in range(12):
        opt.zero_grad()
        loss = e.forward_spectral_ncc(est, ft, view="ap", high_pass=True)
        loss.backward()
        opt.step()
        losses.append(float(loss))
    assert losses[-1] < losses[0]
    # the image-space loss descends too (reference: test_spectral_ncc_adam_optimization)
    est2 = torch.tensor([[2.0, -1.5, 0.0, 3.0, -2.0, 0.0]], requires_grad=True)
    opt2 = torch.optim.Adam([est2], lr=0.15)
    l2 = []
    for _ in range(12):
        opt2.zero_grad()
        loss = e.spectral_ncc_loss(e(est2, view="ap", out_res=32), tgt, high_pass=True)
        loss.backward()
        opt2.step()
        l2.append(float(loss))
    assert l2[-1] < l2[0]


@pytest.mark.parametrize("dtype,tol", [(torch.bfloat16, 2e-2), (torch.float16, 2e-3)])
def test_reduced_precision_storage(dtype, tol):
    from fvr import TorchDiffFVREngine
    vol = _phantom()
    e32 = TorchDiffFVREngine(torch.from_numpy(vol))
    e16 = TorchDiffFVREngine(torch.from_numpy(vol)).to_precision(dtype)
    p = torch.tensor([[1.0, -1.0, 0.5, 2.0, -1.0, 0.0]], requires_grad=True)
    out = e16(p, view="ap", out_res=32)
    assert out.dtype == torch.float32
    out.sum().backward()
    assert bool(torch.all(torch.isfinite(p.grad))) and float(p.grad.abs().sum()) > 1e-4
    with torch.no_grad():
        ft = e32.prepare_spectral_target(e32(torch.zeros((1, 6)), view="ap", out_res=32))
    q = torch.tensor([[2.0, -1.0, 0.5, 0.5, -0.5, 0.0]])
    q1, q2 = q.clone().requires_grad_(True), q.clone().requires_grad_(True)
    L32 = e32.forward_spectral_ncc(q1, ft)
    L16 = e16.forward_spectral_ncc(q2, ft)
    L32.backward()
    L16.backward()
    assert abs(float(L32) - float(L16)) <= tol * max(float(L32), 1e-3)
    assert float((q1.grad - q2.grad).abs().max()) <= 10 * tol * float(q1.grad.abs().max())


@pytest.mark.parametrize("view", ["ap", "lateral"])
def test_device_t_setup_matches_host_setup(eng, view):
    """The device convention-T set-up (fvr_pt_build_items) reproduces the host builder's
    items and Jacobians, and its combine reproduces the host float64 combine."""
    from fvr import _items, _lib
    from fvr.torch_engine import combine_t
    e, S3 = eng
    ev = e._fused()
    R = 24
    ft, _ = _target(e, S3, [1.0, -0.5, 0.5, 0.4, -0.3, 0.0], view, R)
    tg = torch.cat([ft.reshape(-1, R, R)] * 5).to(torch.complex64).contiguous()  # one target per pose
    poses = np.array([[3.0, -2.0, 1.0, 1.5, -1.0, 0.3], [-4.0, 2.5, -1.5, -0.8, 1.2, 0.0], [0.0] * 6,
                      [12.0, -7.0, 9.0, 2.0, 3.0, -1.0], [-0.5, 0.25, 30.0, -2.5, 0.5, 4.0]])
    P = torch.from_numpy(poses).cuda()
    ncc, grad = ev.device_loss_and_grad(P, tg, view, 4.0)
    ncc, grad = ncc.cpu().numpy().copy(), grad.cpu().numpy().copy()
    got = ev.device_items(len(poses))
    tmpl = ev._template(view, R, _lib.ITEM_ADJOINT, 4.0).copy()
    stride = tg.stride(0) * tg.element_size()
    ptrs = tg.data_ptr() + stride * np.arange(len(poses), dtype=np.uint64)
    ref, jac_ref, _ = _items.t_items(tmpl, ev.pad_shape, poses, view, R, ev.spacing0, ev.rows, ptrs)
    for f in _lib.ITEM_DTYPE.names:
        if f in ("c", "ustep", "vstep"):
            assert np.max(np.abs(got[f].astype(np.int64) - ref[f].astype(np.int64))) <= 4, f
        else:
            np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    jac = ev.ws.bufs["t_jac"][: 18 * len(poses)].cpu().numpy().reshape(-1, 2, 3, 3)
    np.testing.assert_allclose(jac, jac_ref, atol=1e-13)
    # host combine of the same moments: reduce the partials per item on the device, combine in float64
    from fvr.sncc import run_items
    mom = run_items(ev.packed, ref, int(ref["flags"][0])).cpu().numpy()
    t = tg.to(torch.complex128)
    e_t = (t.real.square() + t.imag.square()).sum(dim=(-2, -1)).cpu().numpy()
    nr, dr = combine_t(mom, jac_ref, e_t, _items.t_phase_span(R, ev.spacing0))
    np.testing.assert_allclose(ncc, nr, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(grad, -dr / len(poses), rtol=1e-7, atol=1e-12 * np.abs(dr).max())


def test_target_energy_cache_follows_in_place_changes(eng):
    """The fused path keeps sum |T|^2 of the target tensor it last saw: a second call with
    the same tensor gives the same loss bit for bit, and an in-place change of that tensor
    is seen (the loss equals that of a fresh copy of the changed tensor, bit for bit)."""
    e, S3 = eng
    R = 24
    ft, _ = _target(e, S3, [1.0, -0.5, 0.5, 0.4, -0.3, 0.0], "ap", R)
    ft = ft.to(e.device, torch.complex64).contiguous()
    P = torch.tensor([[3.0, -2.0, 1.0, 1.5, -1.0, 0.3], [-4.0, 2.5, -1.5, -0.8, 1.2, 0.0]],
                     dtype=torch.float64, device=e.device)
    l0 = float(e.forward_spectral_ncc(P, ft))
    assert float(e.forward_spectral_ncc(P, ft)) == l0
    ft[..., : R // 3, :] = 0  # in place: the version counter moves
    l1 = float(e.forward_spectral_ncc(P, ft))
    l2 = float(e.forward_spectral_ncc(P, ft.clone()))
    assert l1 == l2 and l1 != l0
