"""GPU parity of the PoU-FVR perspective path (PerspectiveSlabFVR, pou_pinhole_warp_sum)
against the FP64 oracle (oracle/cone_beam_np.py) and the reference's own checks
(tests/test_cone_beam_fvr.py of the reference) on the HIP path.

Tolerances: slab images and projections within 1e-4 of max|ref| (float32
interpolation, float32 LDS FFT); warp outputs 1e-5; gradients 1e-4 of max
against float64 torch autograd of the reference's grid_sample warp, and 2e-3
of max between the fused adjoint and the generic composition.
"""

import numpy as np
import pytest
import torch

from oracle import cone_beam_np as C

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _vol(n=24):
    vol = np.zeros((n, n, n), dtype=np.float32)
    zz, yy, xx = np.ogrid[:n, :n, :n]
    vol[((zz - 11) ** 2 + (yy - 13) ** 2 + (xx - 12) ** 2) < 7 ** 2] = 0.08
    vol[9:15, 5:9, 14:19] += 0.05
    return vol


@pytest.fixture(scope="module")
def model():
    from fvr import PerspectiveSlabFVR
    vol = _vol()
    m = PerspectiveSlabFVR.from_volume(vol, vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=16)
    o = C.from_volume(vol.astype(np.float64), 2.0, (2, 2, 2), 16)
    return m, o


def _lattice(H=20, W=22):
    from fvr import PerspectiveSlabFVR
    k_inv = torch.tensor([[1.0 / 300.0, 0.0, -W / 600.0], [0.0, 1.0 / 300.0, -H / 600.0], [0.0, 0.0, 1.0]])
    lat = PerspectiveSlabFVR.make_pinhole_lattice(k_inv, 600.0, H, W)
    lat_np = C.make_pinhole_lattice(k_inv.double().numpy(), 600.0, H, W)
    return lat, lat_np


def _se3(rotvec, t):
    from scipy.spatial.transform import Rotation
    M = np.tile(np.eye(4), (len(rotvec), 1, 1))
    M[:, :3, :3] = Rotation.from_rotvec(rotvec).as_matrix()
    M[:, :3, 3] = t
    return M


MATS = _se3(np.array([[0.15, -0.10, 0.08], [-0.20, 0.12, -0.05], [0.0, 0.0, 0.0]]),
            np.array([[5.0, -4.0, -300.0], [-3.0, 6.0, -310.0], [0.0, 0.0, -300.0]]))


def test_blocks_match_oracle(model):
    m, o = model
    assert (m.B, m.P) == (o["B"], o["P"]) and abs(m.ds_block - o["ds"]) < 1e-12
    np.testing.assert_allclose(m.centers.cpu().numpy(), o["centers"], atol=1e-4)
    F = m.F_blk_ri.double().cpu().numpy()
    ref = o["F_blk"]
    assert np.max(np.abs(F[:, 0] + 1j * F[:, 1] - ref)) <= 1e-6 * np.abs(ref).max()


def test_fused_slab_images_match_oracle(model):
    m, o = model
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.standard_normal((3, 3, 3)))[0]
    u, v = Q[:, :, 0], Q[:, :, 1]
    assert m.fused_ok()
    imgs = m._slice_and_ifft(torch.from_numpy(u).float(), torch.from_numpy(v).float()).cpu().numpy()
    ref = C.slab_images(o, u, v)
    assert np.max(np.abs(imgs - ref)) <= 1e-4 * np.abs(ref).max()
    m.fused_kernels = False
    try:
        gen = m._slice_and_ifft(torch.from_numpy(u).float(), torch.from_numpy(v).float()).cpu().numpy()
    finally:
        m.fused_kernels = True
    assert np.max(np.abs(gen - ref)) <= 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("P", [4, 8, 16, 32, 64])
def test_half_spectrum_slab_images_match_oracle_at_every_lds_size(P):
    """The fused slice + inverse FFT (Hermitian half spectrum: a sample and its mirror per
    stored entry, column inverse FFTs, the even / odd pre-pass, row inverse FFTs of length
    P/2) against the oracle's full complex transform, for every power-of-two P the LDS kernel
    takes below 128 (128: the full-size test), float32 and bfloat16 block spectra."""
    from fvr import PerspectiveSlabFVR
    n = max(8, P)
    vol = np.zeros((n, n, n), dtype=np.float32)
    zz, yy, xx = np.ogrid[:n, :n, :n]
    vol[((zz - n * 0.45) ** 2 + (yy - n * 0.55) ** 2 + (xx - n * 0.5) ** 2) < (0.3 * n) ** 2] = 0.08
    vol[n // 3:n // 2, n // 5:n // 3, n // 2:2 * n // 3] += 0.05
    o = C.from_volume(vol.astype(np.float64), 2.0, (2, 2, 2), P // 2)
    rng = np.random.default_rng(P)
    Q = np.linalg.qr(rng.standard_normal((4, 3, 3)))[0]
    u, v = Q[:, :, 0], Q[:, :, 1]
    u[0], v[0] = [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]  # axis-aligned: samples on lattice points
    ref = C.slab_images(o, u, v)
    for dt, tol in ((torch.float32, 1e-4), (torch.bfloat16, 2e-2)):
        m = PerspectiveSlabFVR.from_volume(vol, vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=P // 2, dtype=dt)
        assert m.P == P and m.fused_ok()
        imgs = m._slice_and_ifft(torch.from_numpy(u).float(), torch.from_numpy(v).float()).cpu().numpy()
        assert imgs.shape == ref.shape
        err = np.abs(imgs - ref).max() / np.abs(ref).max()
        assert err <= tol, (P, dt, err)


def test_warp_forward_and_gradients_match_grid_sample_autograd():
    from fvr import pou_pinhole_warp_sum
    from test_oracle_pou import _warp_torch_f64, warp_case
    imgs, d_b, cu, cv, rx, ry, ic, inv_ds, c_idx = warp_case()
    ref_in = [t.clone().requires_grad_(True) for t in (imgs, d_b, cu, cv)]
    ref = _warp_torch_f64(*ref_in, rx, ry, ic, inv_ds, c_idx)
    w = torch.randn(ref.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(4))
    (ref * w).sum().backward()
    got_in = [t.float().to(DEV).requires_grad_(True) for t in (imgs, d_b, cu, cv)]
    out = pou_pinhole_warp_sum(*got_in, rx.float().to(DEV), ry.float().to(DEV), ic.float().to(DEV),
                               inv_ds=inv_ds, c_idx=c_idx)
    assert out.shape == ref.shape
    assert float((out.detach().double().cpu() - ref.detach()).abs().max()) <= 1e-5 * float(ref.detach().abs().max())
    (out * w.float().to(DEV)).sum().backward()
    for a, b in zip(got_in, ref_in):
        ga, gb = a.grad.double().cpu(), b.grad
        assert float((ga - gb).abs().max()) <= 1e-4 * float(gb.abs().max()), (ga - gb).abs().max()
    # deterministic: a second backward gives identical bits
    got2 = [t.float().to(DEV).requires_grad_(True) for t in (imgs, d_b, cu, cv)]
    out2 = pou_pinhole_warp_sum(*got2, rx.float().to(DEV), ry.float().to(DEV), ic.float().to(DEV),
                                inv_ds=inv_ds, c_idx=c_idx)
    (out2 * w.float().to(DEV)).sum().backward()
    for a, b in zip(got_in, got2):
        assert torch.equal(a.grad, b.grad)


def test_pou_texels_mask_covers_every_weighted_texel(model):
    """fvr.footprint.pou_texels_mask (the slab-image texels the PoU step's algorithmic byte
    count charges to the warp and the adjoint): zeroing every texel outside it leaves the
    warp's output unchanged bit for bit, so it covers every texel the warp weights; here the
    detector sees part of the volume, so it is a strict, non-empty subset of the images."""
    from fvr import pou_pinhole_warp_sum
    from fvr.footprint import pou_texels_mask, pou_texels_read
    m, _ = model
    lat, _ = _lattice()
    mats = torch.from_numpy(MATS).float().to(DEV)
    mask, (d_b, cu, cv, rx, ry, ic) = pou_texels_mask(m, mats, lat)
    assert mask.shape == (m.B, len(MATS), m.P, m.P)
    assert 0 < int(mask.sum()) < mask.numel()
    assert pou_texels_read(m, mats, lat) == int(mask.sum())
    imgs = torch.randn(mask.shape, device=DEV, generator=torch.Generator(device=DEV).manual_seed(5))
    kw = dict(inv_ds=1.0 / m.ds_block, c_idx=float(m.P // 2))
    full = pou_pinhole_warp_sum(imgs, d_b, cu, cv, rx, ry, ic, **kw)
    kept = pou_pinhole_warp_sum(imgs * mask, d_b, cu, cv, rx, ry, ic, **kw)
    assert torch.equal(full, kept)
    assert not torch.equal(full, pou_pinhole_warp_sum(imgs * ~mask, d_b, cu, cv, rx, ry, ic, **kw))


def test_warp_image_gradient_tiny_and_non_finite_upstream_gradients():
    """The fixed-point scatter of the image gradient keeps its relative precision for tiny
    upstream gradients (1e-30: the scale is capped at 2^126, not infinite) and passes NaN
    through: a pose whose upstream gradient holds a NaN gets a NaN image gradient, the
    other poses are unaffected."""
    from fvr import pou_pinhole_warp_sum
    from test_oracle_pou import warp_case
    imgs, d_b, cu, cv, rx, ry, ic, inv_ds, c_idx = warp_case()
    dev_rays = [t.float().to(DEV) for t in (rx, ry, ic)]

    def img_grad(w):
        x = imgs.float().to(DEV).requires_grad_(True)
        out = pou_pinhole_warp_sum(x, *(t.float().to(DEV) for t in (d_b, cu, cv)), *dev_rays, inv_ds=inv_ds,
                                   c_idx=c_idx)
        (out * w).sum().backward()
        return x.grad.double().cpu()

    w = torch.randn((3, 18, 20), generator=torch.Generator().manual_seed(5)).to(DEV)
    g1 = img_grad(w)
    g_tiny = img_grad(w * 1e-30)
    assert bool(torch.isfinite(g_tiny).all()) and float(g_tiny.abs().max()) > 0
    assert float((g_tiny * 1e30 - g1).abs().max()) <= 1e-5 * float(g1.abs().max())
    w_nan = w.clone()
    w_nan[1, 3, 4] = float("nan")
    g_nan = img_grad(w_nan)
    assert bool(torch.isnan(g_nan[:, 1]).all())
    assert torch.equal(g_nan[:, 0], g1[:, 0]) and torch.equal(g_nan[:, 2], g1[:, 2])


@pytest.mark.parametrize("fused", [True, False])
def test_forward_pinhole_matches_oracle(model, fused):
    m, o = model
    lat, lat_np = _lattice()
    m.fused_kernels = fused
    try:
        out = m.forward_pinhole(torch.from_numpy(MATS).float(), lat).cpu().numpy()
    finally:
        m.fused_kernels = True
    ref = C.forward_pinhole(o, MATS, lat_np)
    assert out.shape == ref.shape == (3, 1, 20, 22)
    assert np.max(np.abs(out - ref)) <= 1e-4 * np.abs(ref).max()


def _loss_grad(m, mats, lat, w):
    M = torch.from_numpy(mats).float().to(DEV).requires_grad_(True)
    out = m.forward_pinhole(M, lat)
    (out * w).sum().backward()
    return M.grad.double().cpu().numpy()


def test_fused_adjoint_matches_generic_composition_and_fd(model):
    m, o = model
    lat, lat_np = _lattice()
    w = torch.randn((3, 1, 20, 22), generator=torch.Generator().manual_seed(9)).to(DEV)
    g_fused = _loss_grad(m, MATS, lat, w)
    m.fused_kernels = False
    try:
        g_gen = _loss_grad(m, MATS, lat, w)
    finally:
        m.fused_kernels = True
    assert np.max(np.abs(g_fused - g_gen)) <= 2e-3 * np.abs(g_gen).max(), np.abs(g_fused - g_gen).max()
    # float64 central differences of the oracle on the translation entries
    wn = w.double().cpu().numpy()
    h = 1e-2
    for n in range(2):
        for k in range(3):
            mp, mm = MATS.copy(), MATS.copy()
            mp[n, k, 3] += h
            mm[n, k, 3] -= h
            fd = (np.sum(C.forward_pinhole(o, mp, lat_np) * wn) - np.sum(C.forward_pinhole(o, mm, lat_np) * wn)) / (2 * h)
            assert abs(fd - g_fused[n, k, 3]) <= 2e-3 * max(abs(fd), np.abs(g_fused[:, :3, 3]).max()), (n, k, fd,
                                                                                                      g_fused[n, k, 3])


# --- the reference's PoU checks (tests/test_cone_beam_fvr.py) on the HIP path ------------------
def test_reference_batched_pinhole_and_half_precision():
    from fvr import PerspectiveSlabFVR
    n = 32
    vol = np.zeros((n, n, n), dtype=np.float32)
    zz, yy, xx = np.ogrid[:n, :n, :n]
    vol[((zz - 14) ** 2 + (yy - 17) ** 2 + (xx - 16) ** 2) < 9 ** 2] = 0.08
    fvr = PerspectiveSlabFVR.from_volume(vol, vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=32)
    H = W = 48
    k_inv = torch.tensor([[[1.0 / 300.0, 0.0, -W / 600.0], [0.0, 1.0 / 300.0, -H / 600.0], [0.0, 0.0, 1.0]]])
    lattice = fvr.make_pinhole_lattice(k_inv, sdd_mm=600.0, height=H, width=W)
    mats = torch.eye(4).unsqueeze(0).repeat(4, 1, 1)
    mats[:, 2, 3] = -300.0
    mats[:, 0, 3] = torch.tensor([0.0, 4.0, -3.0, 6.0])
    mats[:, 1, 3] = torch.tensor([0.0, -2.0, 5.0, 1.0])
    mats.requires_grad_(True)
    batched = fvr.forward_pinhole(mats, lattice)
    assert batched.shape == (4, 1, H, W)
    for b in range(4):
        single = fvr.forward_pinhole(mats[b:b + 1].detach(), lattice)
        assert torch.allclose(batched[b:b + 1], single, atol=1e-4)
    batched.sum().backward()
    assert bool(torch.all(torch.isfinite(mats.grad))) and float(mats.grad.abs().sum()) > 1e-4
    half = PerspectiveSlabFVR.from_volume(vol, vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=32).to_precision(
        torch.bfloat16)
    mh = mats.detach().clone().requires_grad_(True)
    oh = half.forward_pinhole(mh, lattice)
    oh.sum().backward()
    assert oh.dtype == torch.float32
    assert torch.allclose(oh, batched.detach(), atol=0.15, rtol=0.05)
    assert float(mh.grad.abs().sum()) > 1e-4


def _torch_se3(rot_vec, trans):
    """SE(3) from rotation vectors via the matrix exponential (as the reference's test builds it)."""
    skew = torch.zeros((rot_vec.shape[0], 3, 3), dtype=rot_vec.dtype)
    skew[:, 0, 1], skew[:, 1, 0] = -rot_vec[:, 2], rot_vec[:, 2]
    skew[:, 0, 2], skew[:, 2, 0] = rot_vec[:, 1], -rot_vec[:, 1]
    skew[:, 1, 2], skew[:, 2, 1] = -rot_vec[:, 0], rot_vec[:, 0]
    mats = torch.eye(4, dtype=rot_vec.dtype).unsqueeze(0).repeat(rot_vec.shape[0], 1, 1)
    mats[:, :3, :3] = torch.matrix_exp(skew)
    mats[:, :3, 3] = trans
    return mats


def test_reference_invariant_matches_explicit_ray_path(model):
    """SE(3)-invariant lattice path == explicit per-pose rays, outputs and 6-DoF gradients
    (tests/test_cone_beam_fvr.py:230-274 of the reference, on-manifold parameters)."""
    m, _ = model
    lat, _ = _lattice(28, 28)
    rot0 = torch.tensor([[0.15, -0.10, 0.08], [-0.20, 0.12, -0.05]])
    xyz0 = torch.tensor([[5.0, -4.0, -300.0], [-3.0, 6.0, -310.0]])
    r_a, t_a = rot0.clone().requires_grad_(True), xyz0.clone().requires_grad_(True)
    r_b, t_b = rot0.clone().requires_grad_(True), xyz0.clone().requires_grad_(True)
    out_inv = m.forward_pinhole(_torch_se3(r_a, t_a), lat)
    out_exp = m.forward_pinhole(_torch_se3(r_b, t_b), lat[:5])
    w = torch.randn(out_inv.shape, generator=torch.Generator().manual_seed(2)).to(DEV)
    (out_inv * w).sum().backward()
    (out_exp * w).sum().backward()
    assert float(out_inv.abs().max()) > 0
    assert torch.allclose(out_inv, out_exp, atol=1e-4, rtol=1e-4)
    assert torch.allclose(r_a.grad, r_b.grad, atol=5e-4, rtol=1e-3), (r_a.grad - r_b.grad).abs().max()
    assert torch.allclose(t_a.grad, t_b.grad, atol=5e-4, rtol=1e-3), (t_a.grad - t_b.grad).abs().max()


def test_reference_carm_gradients_and_magnification():
    from fvr import PerspectiveSlabFVR
    n = 48
    vol = np.zeros((n, n, n), dtype=np.float32)
    zz, yy, xx = np.ogrid[:n, :n, :n]
    vol[((zz - 24) ** 2 + (yy - 24) ** 2 + (xx - 24) ** 2) < 12 ** 2] = 0.05
    cb = PerspectiveSlabFVR.from_volume(vol, vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=40)  # P = 80: generic path
    t = [torch.tensor(x) for x in (3.0, -2.0, 1.5, 1.0, -1.0, 0.5)]
    target = cb.forward_carm_view(*t, sad_mm=600.0, sdd_mm=900.0, det_px=48, det_fov_mm=96.0).detach()
    params = [torch.tensor(0.0, requires_grad=True) for _ in range(6)]
    pred = cb.forward_carm_view(*params, sad_mm=600.0, sdd_mm=900.0, det_px=48, det_fov_mm=96.0)
    a0, b0 = pred - pred.mean(), target - target.mean()
    loss = 1.0 - (a0 * b0).sum() / (a0.norm() * b0.norm() + 1e-8)
    loss.backward()
    assert all(p.grad is not None and bool(torch.isfinite(p.grad)) for p in params)
    assert any(abs(float(p.grad)) > 1e-4 for p in params)
    # front/back magnification (reference sizes)
    n, vox = 64, 2.0
    vol = np.zeros((n, n, n), dtype=np.float32)
    c = (np.arange(n) - (n - 1) / 2.0) * vox
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    vol[((zz - 16.0) ** 2 + (yy + 32.0) ** 2 + (xx - 24.0) ** 2) < 9.0 ** 2] = 1.0
    vol[((zz + 16.0) ** 2 + (yy - 32.0) ** 2 + (xx - 24.0) ** 2) < 9.0 ** 2] = 1.0
    f = PerspectiveSlabFVR.from_volume(vol, vox_mm=vox, grid_blocks=(2, 4, 2), block_res=48)
    z0 = torch.tensor(0.0)
    img = f.forward_carm_view(z0, z0, z0, z0, z0, z0, sad_mm=300.0, sdd_mm=600.0, det_px=96, det_fov_mm=n * vox)
    u = torch.linspace(-0.5, 0.5, 96, device=img.device) * (n * vox)
    top, bot = img[48:, :], img[:48, :]
    x_front = float((top.sum(0) * u).sum() / top.sum())
    x_back = float((bot.sum(0) * u).sum() / bot.sum())
    assert x_front > x_back + 7.0


@pytest.mark.parametrize("view", ["ap", "lateral"])
@pytest.mark.parametrize("block_res", [32, 40], ids=["fused-P64", "generic-P80"])
def test_carm_view_and_pose_gradient_match_oracle(view, block_res):
    """forward_carm_view (cone_beam_fvr.py:761-832: Euler pose in degrees and voxels, AP or
    lateral C-arm geometry) against the oracle's restatement, and its 6-DoF gradient
    (autograd through the explicit-ray path: the fused or the generic kernels and the ray
    gradients) against FP64 central differences of the oracle (h = 1e-5 deg / voxel: the
    bilinear warp has kinks at pixel boundaries, and differences over h = 1e-3 straddle
    enough of them to be off by 2e-3; at 1e-5 they have converged to 1e-7):
    images within 1e-4 of max, gradient within 1e-4 of max|grad|."""
    from fvr import PerspectiveSlabFVR
    n, vox = 32, 2.0
    vol = np.zeros((n, n, n), dtype=np.float32)
    zz, yy, xx = np.ogrid[:n, :n, :n]
    vol[((zz - 14) ** 2 + (yy - 17) ** 2 + (xx - 15) ** 2) < 8 ** 2] = 0.06
    vol[8:14, 18:25, 9:15] += 0.04
    m = PerspectiveSlabFVR.from_volume(vol, vox_mm=vox, grid_blocks=(2, 2, 2), block_res=block_res)
    o = C.from_volume(vol.astype(np.float64), vox, (2, 2, 2), block_res)
    geo = dict(sad_mm=600.0, sdd_mm=900.0, det_px=40, det_fov_mm=96.0, view=view)
    x0 = np.array([3.0, -2.0, 1.5, 1.0, -1.0, 0.5])
    params = [torch.tensor(float(v), requires_grad=True) for v in x0]
    img = m.forward_carm_view(*params, **geo)
    ref = C.forward_carm_view(o, *x0, **geo)
    assert img.shape == ref.shape == (40, 40)
    assert np.abs(img.detach().double().cpu().numpy() - ref).max() <= 1e-4 * np.abs(ref).max()
    w = np.random.default_rng(4).standard_normal((40, 40))
    (img * torch.from_numpy(w).float().to(img.device)).sum().backward()
    g = np.array([float(p.grad) for p in params])
    h = 1e-5
    fd = np.array([(np.sum(C.forward_carm_view(o, *(x0 + h * e), **geo) * w)
                    - np.sum(C.forward_carm_view(o, *(x0 - h * e), **geo) * w)) / (2 * h) for e in np.eye(6)])
    assert np.abs(g - fd).max() <= 1e-4 * np.abs(fd).max(), (g, fd)


def test_reference_presets_and_offsets():
    from fvr import SLAB_PRESETS, PerspectiveSlabFVR, sample_isocenter_parity_offsets
    assert SLAB_PRESETS["isotropic"]["grid_blocks"] == (4, 4, 4) and SLAB_PRESETS["isotropic"]["block_res"] == 64
    vol = np.zeros((32, 32, 32), dtype=np.float32)
    vol[8:24, 8:24, 8:24] = 0.5
    iso = PerspectiveSlabFVR.from_volume(vol, vox_mm=2.0, preset="isotropic", block_res=16)
    assert iso.B == 64 and iso.F_blk_ri.dtype == torch.bfloat16
    with pytest.raises(ValueError):
        PerspectiveSlabFVR.resolve_slab_preset("nonexistent")
    offs = sample_isocenter_parity_offsets((1.0, 2.0, 4.0))
    assert offs.shape == (27, 3)
    assert torch.allclose(offs.abs().max(dim=0).values, torch.tensor([0.5, 1.0, 2.0]))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.bfloat16])
def test_packed_block_layout_image(dtype):
    """The fused kernels' gather layout, read back and decoded by the header's definition:
    per block (P+3) x (P+2) x (P+2) elements, storage index = logical + 1, zero margins;
    element (z, y, x) = (re, im) as float2, or for 16-bit spectra the (re, im) words (16-bit re
    in the low half, im in the high half) of rows y and y + 1 (row P + 2 zero)."""
    from fvr import _lib
    L = _lib.lib()
    dev = torch.device("cuda", 0)
    B, P = 3, 8
    g = torch.Generator().manual_seed(3)
    F = torch.randn((B, 2, P, P, P), generator=g).to(dtype).to(dev)
    code = {torch.float32: 0, torch.float16: 1, torch.bfloat16: 2}[dtype]
    nb = int(L.fvr_pou_packed_bytes(code, B, P))
    assert nb == B * (P + 3) * (P + 2) * (P + 2) * 8
    buf = torch.full((nb,), 0x5A, dtype=torch.uint8, device=dev)  # every byte must be written
    _lib.check(L.fvr_pou_pack_blocks(F.data_ptr(), code, B, P, buf.data_ptr(), _lib.stream_ptr(dev)), "pack")
    torch.cuda.synchronize()
    raw = buf.cpu()
    ent = torch.zeros((B, P + 3, P + 3, P + 2, 2), dtype=dtype)  # one extra zero row
    Fc = F.cpu()
    ent[:, 1:P + 1, 1:P + 1, 1:P + 1, 0] = Fc[:, 0]
    ent[:, 1:P + 1, 1:P + 1, 1:P + 1, 1] = Fc[:, 1]
    if code == 0:
        got = raw.view(torch.float32).view(B, P + 3, P + 2, P + 2, 2)
        assert torch.equal(got, ent[:, :, :P + 2])
    else:
        got = raw.view(dtype).view(B, P + 3, P + 2, P + 2, 2, 2)
        assert torch.equal(got[..., 0, :], ent[:, :, :P + 2])
        assert torch.equal(got[..., 1, :], ent[:, :, 1:P + 3])


@pytest.mark.parametrize("block_res,H,W", [(8, 20, 22), (16, 20, 22), (32, 64, 48), (64, 96, 80)])
def test_separable_gather_adjoint_matches_the_scan(block_res, H, W):
    """The fused backward of a zero-skew pinhole lattice takes the gather form of the warp
    adjoint (fvr_pou_render_bwd_sep, P = 2 block_res: every LDS size); its pose gradient
    equals the fixed-point scan's (fvr_pou_render_bwd) to float32 summation order, and is
    bit-identical run to run."""
    from fvr import PerspectiveSlabFVR
    m = PerspectiveSlabFVR.from_volume(_vol(), vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=block_res)
    f = 300.0 * H / 20.0
    k_inv = torch.tensor([[1.0 / f, 0.0, -W / (2 * f)], [0.0, 1.0 / f, -H / (2 * f)], [0.0, 0.0, 1.0]])
    lat = [t.to(DEV) for t in PerspectiveSlabFVR.make_pinhole_lattice(k_inv, 600.0, H, W)]
    w = torch.randn((3, 1, H, W), generator=torch.Generator().manual_seed(11)).to(DEV)
    g_sep = _loss_grad(m, MATS, lat, w)
    assert m._sep_rays[3] is not None  # the gather kernel ran
    assert np.array_equal(g_sep, _loss_grad(m, MATS, lat, w))
    m.separable_adjoint = False
    try:
        g_scan = _loss_grad(m, MATS, lat, w)
    finally:
        m.separable_adjoint = True
    assert np.max(np.abs(g_sep - g_scan)) <= 1e-5 * np.abs(g_scan).max(), np.abs(g_sep - g_scan).max()


def test_separable_gather_adjoint_propagates_non_finite_upstream_gradients(model):
    """A NaN in one pose's upstream gradient makes that pose's gradient NaN (the flag of
    fvr_pou_gprep) and leaves the other poses' gradients bit for bit unchanged."""
    m, _ = model
    lat = [t.to(DEV) for t in _lattice()[0]]
    w = torch.randn((3, 1, 20, 22), generator=torch.Generator().manual_seed(12)).to(DEV)
    g1 = _loss_grad(m, MATS, lat, w)
    w_nan = w.clone()
    w_nan[1, 0, 3, 4] = float("nan")
    g2 = _loss_grad(m, MATS, lat, w_nan)
    assert m._sep_rays[3] is not None
    assert np.isnan(g2[1, :3, :]).all()
    assert np.array_equal(g2[0], g1[0]) and np.array_equal(g2[2], g1[2])


def test_fused_pinhole_geometry_matches_the_torch_chain(model):
    """forward_pinhole on a lattice with a ray cache computes the pose geometry (source,
    detector axes, block offsets) in one launch and its adjoint in another
    (fvr_pou_pose_geom / _bwd); projections and pose-matrix gradients equal those of the
    reference's torch composition (cone_beam_fvr.py:670-688, 507-521) over the same
    render operator, to float32 rounding."""
    m, _ = model
    lat = [t.to(DEV) for t in _lattice()[0]]
    w = torch.randn((3, 1, 20, 22), generator=torch.Generator().manual_seed(13)).to(DEV)
    M = torch.from_numpy(MATS).float().to(DEV).requires_grad_(True)
    out = m.forward_pinhole(M, lat)
    (out * w).sum().backward()
    M2 = torch.from_numpy(MATS).float().to(DEV).requires_grad_(True)
    R, t = M2[:, :3, :3], M2[:, :3, 3]
    srcs = torch.einsum("nij,j->ni", R, lat[0]) + t
    u_det, v_det, n_0 = (torch.einsum("nij,j->ni", R, a) for a in lat[2:5])
    ref = m._project_core_invariant(srcs, u_det, v_det, n_0, lat[5:11]).unsqueeze(1)
    (ref * w).sum().backward()
    assert float((out - ref).abs().max()) <= 1e-5 * float(ref.abs().max())
    g, gr = M.grad.double().cpu().numpy(), M2.grad.double().cpu().numpy()
    assert np.max(np.abs(g - gr)) <= 1e-5 * np.abs(gr).max(), np.abs(g - gr).max()
    assert np.all(g[:, 3, :] == 0.0)


@pytest.mark.parametrize("block_res,H,W", [(8, 20, 22), (16, 40, 36), (64, 96, 80), (64, 256, 256)])
def test_tiled_separable_warp_equals_the_per_pixel_warp(block_res, H, W):
    """A zero-skew pinhole lattice takes the tiled forward warp (fvr_pou_warp_fwd_sep: per
    (pose, 32 x 32 tile) the slab-image patch of each block staged in LDS); its projections
    equal the per-pixel warp's (fvr_pou_warp_fwd) bit for bit, ragged tiles included."""
    from fvr import PerspectiveSlabFVR
    m = PerspectiveSlabFVR.from_volume(_vol(), vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=block_res)
    f = 300.0 * H / 20.0
    k_inv = torch.tensor([[1.0 / f, 0.0, -W / (2 * f)], [0.0, 1.0 / f, -H / (2 * f)], [0.0, 0.0, 1.0]])
    lat = [t.to(DEV) for t in PerspectiveSlabFVR.make_pinhole_lattice(k_inv, 600.0, H, W)]
    M = torch.from_numpy(MATS).float().to(DEV)
    with torch.no_grad():
        tiled = m.forward_pinhole(M, lat)
        assert m._sep_rays[3] is not None and m._sep_rays[3][2]
        m.separable_warp = False
        try:
            plain = m.forward_pinhole(M, lat)
        finally:
            m.separable_warp = True
    assert float(plain.abs().max()) > 0
    assert torch.equal(tiled, plain)


def test_fused_render_saves_its_inputs_for_backward(model):
    """The fused operators keep what their backward reads through save_for_backward: changing
    the pose matrices in place after the forward raises autograd's version error instead of
    giving a gradient of other poses."""
    m, _ = model
    lat = [t.to(DEV) for t in _lattice()[0]]
    M = torch.from_numpy(MATS).float().to(DEV).requires_grad_(True)
    P = M * 1.0  # a float32 device tensor the fused geometry reads without a copy
    out = m.forward_pinhole(P, lat)
    with torch.no_grad():
        P[0, 0, 3] += 1.0
    with pytest.raises(RuntimeError, match="modified by an inplace operation"):
        out.sum().backward()


def test_separable_adjoint_orders_give_the_same_gradient(model):
    """The gather-form adjoint's workgroup orders (block-major, pose-major) compute every
    (block, pose) the same way: the pose gradients are equal bit for bit."""
    m, _ = model
    lat = [t.to(DEV) for t in _lattice()[0]]
    w = torch.randn((3, 1, 20, 22), generator=torch.Generator().manual_seed(14)).to(DEV)
    g0 = _loss_grad(m, MATS, lat, w)
    m.sep_adjoint_order = 1
    try:
        g1 = _loss_grad(m, MATS, lat, w)
    finally:
        m.sep_adjoint_order = 0
    assert np.array_equal(g0, g1)


def test_mirror_queue_and_per_thread_pass_agree(model):
    """The gathers' two ways of taking the samples whose DFT mirror is gathered explicitly --
    the compacted per-workgroup queue (the default) and each thread's pass over its own (what a
    workgroup does when they overflow the queue) -- give the same projections bit for bit and
    the same pose gradient up to float rounding (another summation order)."""
    from fvr import _lib
    m, _ = model
    lat = [t.to(DEV) for t in _lattice()[0]]
    w = torch.randn((3, 1, 20, 22), generator=torch.Generator().manual_seed(15)).to(DEV)
    M = torch.from_numpy(MATS).float().to(DEV)
    res = []
    for on in (1, 0):
        prev = _lib.lib().fvr_pou_set_mirror_queue(on)
        try:
            with torch.no_grad():
                out = m.forward_pinhole(M, lat).cpu().numpy()
            res.append((out, _loss_grad(m, MATS, lat, w)))
        finally:
            _lib.lib().fvr_pou_set_mirror_queue(prev)
    assert np.array_equal(res[0][0], res[1][0])
    assert np.abs(res[0][1] - res[1][1]).max() <= 1e-5 * np.abs(res[1][1]).max()
