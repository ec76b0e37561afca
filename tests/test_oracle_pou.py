"""Pins of the PoU-FVR oracle (oracle/cone_beam_np.py).  CPU only."""

import numpy as np
import pytest
import torch

from oracle import cone_beam_np as C


def test_pou_windows_sum_to_unity():
    # the reference's own cases (tests/test_cone_beam_fvr.py:31-35)
    for length, k in [(64, 2), (96, 3), (128, 4)]:
        ws, _ = C.make_pou_1d(length, k, 0.25)
        np.testing.assert_allclose(np.sum(ws, axis=0), 1.0, atol=1e-12)


def test_resample_matches_torch_interpolate():
    rng = np.random.default_rng(0)
    for m, r in ((13, 8), (9, 16), (97, 64)):
        cube = rng.standard_normal((m, m, m))
        ref = torch.nn.functional.interpolate(torch.tensor(cube)[None, None], size=(r, r, r), mode="trilinear",
                                              align_corners=False)[0, 0].numpy()
        np.testing.assert_allclose(C.resample_linear(cube, r), ref, atol=1e-12)


def test_block_bookkeeping_known_answers():
    vol = np.ones((32, 32, 32))
    a = C.from_volume(vol, 2.5, (2, 2, 2), 32, center_mode="pixel_center")
    b = C.from_volume(vol, 2.5, (2, 2, 2), 32, center_mode="grid_center")
    np.testing.assert_allclose(a["centers"] - b["centers"], 1.25, atol=1e-12)  # test_grid_center_vs_pixel_center
    with pytest.raises(ValueError):
        C.from_volume(vol, 1.0, center_mode="invalid")
    v2 = np.zeros((32, 32, 32))
    v2[8:24, 8:24, 8:24] = 0.5
    assert C.from_volume(v2, 2.0, preset="isotropic", block_res=16)["B"] == 64  # test_slab_presets


def test_slab_image_is_the_ray_sum_at_an_axis_aligned_detector():
    rng = np.random.default_rng(1)
    vol = rng.uniform(0, 1, (12, 12, 12))
    m = C.from_volume(vol, 1.5, (2, 2, 2), 8)
    u = np.array([[0.0, 0.0, 1.0]])  # world (axis0, axis1, axis2): u along x, v along y
    v = np.array([[0.0, 1.0, 0.0]])
    imgs = C.slab_images(m, u, v)
    np.testing.assert_allclose(imgs[:, 0], m["cubes"].sum(axis=1), atol=1e-10)


def _warp_torch_f64(imgs, d_b, cu, cv, rx, ry, ic, inv_ds, c_idx):
    """The reference's grid_sample formulation (triton_slice.py:718-745) in float64."""
    B, N, P, _ = imgs.shape
    H, W = rx.shape
    gu = (d_b.view(B, N, 1, 1) * rx.view(1, 1, H, W) - cu.view(B, N, 1, 1)) * inv_ds + c_idx
    gv = (d_b.view(B, N, 1, 1) * ry.view(1, 1, H, W) - cv.view(B, N, 1, 1)) * inv_ds + c_idx
    grid = torch.stack([gu * 2.0 / (P - 1) - 1.0, gv * 2.0 / (P - 1) - 1.0], dim=-1).view(B * N, H, W, 2)
    s = torch.nn.functional.grid_sample(imgs.reshape(B * N, 1, P, P), grid, mode="bilinear", padding_mode="zeros",
                                        align_corners=True).view(B, N, H, W)
    return s.sum(0) * ic.view(1, H, W)


def warp_case(seed=17):
    """The reference's warp-parity case (tests/test_cone_beam_fvr.py:207-228), float64."""
    g = torch.Generator().manual_seed(seed)
    B, N, P, H, W = 4, 3, 24, 18, 20
    imgs = torch.randn(B, N, P, P, generator=g, dtype=torch.float64)
    d_b = 300.0 + torch.randn(B, N, generator=g, dtype=torch.float64) * 15.0
    cu = torch.randn(B, N, generator=g, dtype=torch.float64) * 4.0
    cv = torch.randn(B, N, generator=g, dtype=torch.float64) * 4.0
    rx = torch.linspace(-0.08, 0.08, W, dtype=torch.float64).view(1, W).expand(H, W).contiguous()
    ry = torch.linspace(-0.07, 0.07, H, dtype=torch.float64).view(H, 1).expand(H, W).contiguous()
    ic = 1.0 / torch.sqrt(1.0 + rx ** 2 + ry ** 2).clamp(min=0.1)
    return imgs, d_b, cu, cv, rx, ry, ic, 0.25, 12.0


def test_warp_stencil_equals_grid_sample():
    imgs, d_b, cu, cv, rx, ry, ic, inv_ds, c_idx = warp_case()
    ref = _warp_torch_f64(imgs, d_b, cu, cv, rx, ry, ic, inv_ds, c_idx).numpy()
    got = C.warp_sum(imgs.numpy(), d_b.numpy(), cu.numpy(), cv.numpy(), rx.numpy(), ry.numpy(), ic.numpy(), inv_ds,
                     c_idx)
    np.testing.assert_allclose(got, ref, atol=1e-11)


def _model_and_lattice(n=24, block_res=16, H=20, W=22):
    vol = np.zeros((n, n, n))
    zz, yy, xx = np.ogrid[:n, :n, :n]
    vol[((zz - 11) ** 2 + (yy - 13) ** 2 + (xx - 12) ** 2) < 7 ** 2] = 0.08
    vol[9:15, 5:9, 14:19] += 0.05
    m = C.from_volume(vol, 2.0, (2, 2, 2), block_res)
    k_inv = np.array([[1.0 / 300.0, 0.0, -W / 600.0], [0.0, 1.0 / 300.0, -H / 600.0], [0.0, 0.0, 1.0]])
    lat = C.make_pinhole_lattice(k_inv, 600.0, H, W)
    return vol, m, lat


def _se3(rotvec, t):
    from scipy.spatial.transform import Rotation
    M = np.tile(np.eye(4), (len(rotvec), 1, 1))
    M[:, :3, :3] = Rotation.from_rotvec(rotvec).as_matrix()
    M[:, :3, 3] = t
    return M


def test_invariant_lattice_equals_explicit_rays_and_batched_equals_single():
    _, m, lat = _model_and_lattice()
    mats = _se3(np.array([[0.15, -0.10, 0.08], [-0.20, 0.12, -0.05]]),
                np.array([[5.0, -4.0, -300.0], [-3.0, 6.0, -310.0]]))
    inv = C.forward_pinhole(m, mats, lat)[:, 0]
    R, t = mats[:, :3, :3], mats[:, :3, 3]
    srcs = np.einsum("nij,j->ni", R, lat["src0"]) + t
    trg = np.einsum("nij,hwj->nhwi", R, lat["trg0"]) + t[:, None, None, :]
    exp = C.project_rays_batched(m, srcs, trg, np.einsum("nij,j->ni", R, lat["u0"]),
                                 np.einsum("nij,j->ni", R, lat["v0"]), np.einsum("nij,j->ni", R, lat["n0"]))
    np.testing.assert_allclose(inv, exp, atol=1e-9 * np.abs(exp).max())
    for b in range(2):
        np.testing.assert_allclose(C.forward_pinhole(m, mats[b:b + 1], lat)[0, 0], inv[b], atol=1e-12)


def test_perspective_magnification_front_vs_back():
    # the reference's physics check and sizes (tests/test_cone_beam_fvr.py:37-62)
    n, vox = 64, 2.0
    vol = np.zeros((n, n, n))
    c = (np.arange(n) - (n - 1) / 2.0) * vox
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    vol[((zz - 16.0) ** 2 + (yy + 32.0) ** 2 + (xx - 24.0) ** 2) < 9.0 ** 2] = 1.0
    vol[((zz + 16.0) ** 2 + (yy - 32.0) ** 2 + (xx - 24.0) ** 2) < 9.0 ** 2] = 1.0
    m = C.from_volume(vol, vox, (2, 4, 2), 48)
    img = C.forward_carm_view(m, 0, 0, 0, 0, 0, 0, sad_mm=300.0, sdd_mm=600.0, det_px=96, det_fov_mm=n * vox)
    u = np.linspace(-0.5, 0.5, 96) * n * vox
    top, bot = img[48:, :], img[:48, :]
    x_front = float((top.sum(0) * u).sum() / top.sum())
    x_back = float((bot.sum(0) * u).sum() / bot.sum())
    assert x_front > x_back + 7.0
