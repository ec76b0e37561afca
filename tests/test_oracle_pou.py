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


# ---------------------------------------------------------------------------
# block geometry: every block cube, sampled from world coordinates
# ---------------------------------------------------------------------------
def _blob_volume(shape, vox, spec):
    """Sum of Gaussians (centre mm, sigma mm, amplitude) on a pixel-centred grid."""
    axes = [(np.arange(n) - (n - 1) / 2.0) * vox for n in shape]
    zz, yy, xx = np.meshgrid(*axes, indexing="ij")
    v = np.zeros(shape)
    for (z, y, x), s, a in spec:
        v += a * np.exp(-((zz - z) ** 2 + (yy - y) ** 2 + (xx - x) ** 2) / (2 * s * s))
    return v


BLOBS = [((6.0, -8.0, 4.0), 6.0, 0.02), ((-10.0, 12.0, -6.0), 5.0, 0.03), ((14.0, 10.0, 16.0), 7.0, 0.015)]


def _expected_cube(vol, windows, center, ds, P, block_res, vox, origin, trilinear):
    """The padded cube a block should hold, formed from world coordinates alone: sample
    q of the P^3 cube sits at world point center + (q - P//2) ds; inside the block_res^3
    core it holds ds x sinc^2 pre-compensation x (PoU-weighted volume, trilinear at that
    point); outside the core it is 0.  Returns (expected, volume index coordinates)."""
    pad_b = (P - block_res) // 2
    q = np.arange(pad_b, pad_b + block_res)
    idx = [(center[k] + (q - P // 2) * ds - origin[k]) / vox for k in range(3)]
    g = vol * windows
    pts = np.stack(np.meshgrid(*idx, indexing="ij"), axis=0).reshape(3, -1)
    core = trilinear(g, pts).reshape((block_res,) * 3)
    w = C.sinc2_precomp_1d(block_res, P)
    out = np.zeros((P, P, P))
    sl = slice(pad_b, pad_b + block_res)
    out[sl, sl, sl] = core * ds * (w[:, None, None] * w[None, :, None] * w[None, None, :])
    return out, idx


@pytest.mark.parametrize("block_res", [None, 16, 40], ids=["no-resample", "down", "up"])
def test_block_cubes_equal_world_space_sampling_of_the_weighted_volume(block_res):
    """Pins the block bookkeeping of cone_beam_fvr.py:208-284 (block centres with the
    half-voxel offset to the FFT centre, ds = M vox / block_res, the cube size M, the
    resampling rule, the sinc^2 window's centring, the ds factor, the padding) without
    retyping it: each cube equals the PoU-weighted volume sampled (scipy
    map_coordinates, trilinear) at the world points its samples stand for, wherever the
    trilinear stencil lies inside the block's support (the cube is cut to the support
    where the window exceeds 1e-6); on a non-cubic volume with a non-cubic block grid, so
    a transposed axis fails.  Without resampling every sample lands on a voxel centre.
    The same comparison with the centres moved by half a voxel fails (the check is
    sensitive to the plausible mistake)."""
    from scipy.ndimage import map_coordinates

    def trilinear(g, pts):
        return map_coordinates(g, pts, order=1, mode="constant", cval=0.0)

    shape, vox, grid = (36, 40, 44), 2.0, (2, 3, 2)
    vol = _blob_volume(shape, vox, BLOBS)
    wl = [C.make_pou_1d(n, k, 0.25)[0] for n, k in zip(shape, grid)]
    probe = C.from_volume(vol, vox, grid, 8, energy_thresh=-1.0)
    M = int(round(probe["ds"] * 8 / vox))  # the cube size the bookkeeping chose
    br = M if block_res is None else block_res
    m = C.from_volume(vol, vox, grid, br, energy_thresh=-1.0)
    assert m["B"] == 12 and m["P"] == 2 * br
    assert abs(m["ds"] - M * vox / br) < 1e-15
    origin = [-0.5 * (n - 1) * vox for n in shape]
    b = 0
    worst, worst_shift = 0.0, 0.0
    for k0 in range(grid[0]):
        for k1 in range(grid[1]):
            for k2 in range(grid[2]):
                win = wl[0][k0][:, None, None] * wl[1][k1][None, :, None] * wl[2][k2][None, None, :]
                sup = [np.nonzero(w[k] > 1e-6)[0][[0, -1]] for w, k in zip(wl, (k0, k1, k2))]
                exp, idx = _expected_cube(vol, win, m["centers"][b], m["ds"], m["P"], br, vox, origin, trilinear)
                pad_b = (m["P"] - br) // 2
                inside = [(x >= lo + 1) & (x <= hi - 1) for x, (lo, hi) in zip(idx, sup)]
                mask = np.zeros((m["P"],) * 3, bool)
                sl = slice(pad_b, pad_b + br)
                mask[sl, sl, sl] = inside[0][:, None, None] & inside[1][None, :, None] & inside[2][None, None, :]
                assert mask.sum() > 0.05 * br ** 3  # not vacuous: corner blocks have short supports
                worst = max(worst, np.abs(m["cubes"][b] - exp)[mask].max())
                exp_s, _ = _expected_cube(vol, win, m["centers"][b] + 0.5 * vox, m["ds"], m["P"], br, vox, origin,
                                          trilinear)
                worst_shift = max(worst_shift, np.abs(m["cubes"][b] - exp_s)[mask].max())
                b += 1
    peak = np.abs(m["cubes"]).max()
    assert worst <= 1e-12 * peak, worst / peak
    assert worst_shift >= 1e-2 * peak, worst_shift / peak


def test_block_cubes_re_sum_to_the_volume():
    """Without resampling, the cubes with their sinc^2 window and ds factor divided out,
    placed back at the voxels their world coordinates name, sum to the volume (the
    partition of unity, up to the 1e-6 support cut)."""
    shape, vox, grid = (36, 40, 44), 2.0, (2, 3, 2)
    vol = _blob_volume(shape, vox, BLOBS)
    probe = C.from_volume(vol, vox, grid, 8, energy_thresh=-1.0)
    M = int(round(probe["ds"] * 8 / vox))
    m = C.from_volume(vol, vox, grid, M, energy_thresh=-1.0)
    P, ds = m["P"], m["ds"]
    assert abs(ds - vox) < 1e-15
    pad_b = (P - M) // 2
    w = C.sinc2_precomp_1d(M, P)
    w3 = w[:, None, None] * w[None, :, None] * w[None, None, :]
    origin = np.array([-0.5 * (n - 1) * vox for n in shape])
    acc = np.zeros(shape)
    for b in range(m["B"]):
        core = m["cubes"][b][pad_b:pad_b + M, pad_b:pad_b + M, pad_b:pad_b + M] / (w3 * ds)
        start = (m["centers"][b] + (pad_b - P // 2) * ds - origin) / vox  # volume index of core sample 0
        assert np.allclose(start, np.round(start), atol=1e-9)
        s = np.round(start).astype(int)
        lo = np.maximum(s, 0)
        hi = np.minimum(s + M, shape)
        acc[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += core[lo[0] - s[0]:hi[0] - s[0], lo[1] - s[1]:hi[1] - s[1],
                                                          lo[2] - s[2]:hi[2] - s[2]]
    assert np.abs(acc - vol).max() <= 1e-5 * np.abs(vol).max()


def test_sinc2_precomp_closed_form():
    """1 / max(sinc^2((i - (n-1)/2) / p), 0.05) (cone_beam_fvr.py:64-68) from sin(pi c)/(pi c)."""
    import math
    for n, p in ((5, 10), (8, 16), (64, 128), (7, 7)):
        got = C.sinc2_precomp_1d(n, p)
        for i in range(n):
            c = (i - (n - 1) / 2.0) / p
            s = 1.0 if c == 0 else math.sin(math.pi * c) / (math.pi * c)
            assert abs(got[i] - 1.0 / max(s * s, 0.05)) <= 1e-13 * got[i]
        np.testing.assert_allclose(got, got[::-1], rtol=1e-15)
    assert C.sinc2_precomp_1d(5, 10)[2] == 1.0  # odd n: the centre sample is c = 0
    assert C.sinc2_precomp_1d(3, 1)[0] == 20.0  # |c| = 1: sinc = 0, the clamp at 0.05


def test_pinhole_projection_against_brute_force_ray_sums():
    """End-to-end geometry: the PoU projection of a smooth phantom against brute-force
    line integrals (trilinear, 0.1 mm steps) along the true pinhole rays, source 3 m from
    the isocentre so that each block's single magnification is close to exact.  What is
    left is the method's own approximation (the sinc^2 pre-compensation over-weights the
    low frequencies by a few percent): image centroids agree within 0.3 px (measured
    0.07-0.19 px) and values within 12 % of the peak.  A block centre moved by 1 mm (half
    a voxel) moves the centroid by 0.4-0.65 px and fails."""
    from scipy.ndimage import map_coordinates
    from scipy.spatial.transform import Rotation
    n, vox = 40, 2.0
    vol = _blob_volume((n, n, n), vox, BLOBS)
    sad, sdd, H, pix = 3000.0, 6000.0, 64, 3.0
    f = sdd / pix
    k_inv = np.array([[1 / f, 0, -H / (2 * f)], [0, 1 / f, -H / (2 * f)], [0, 0, 1]])
    lat = C.make_pinhole_lattice(k_inv, sdd, H, H)
    rots = Rotation.from_euler("zyx", [[0, 0, 0], [20, -15, 10], [-40, 30, 60]], degrees=True).as_matrix()
    mats = np.tile(np.eye(4), (3, 1, 1))
    mats[:, :3, :3] = rots
    mats[:, :3, 3] = np.einsum("nij,j->ni", rots, [0, 0, -sad]) + [[3.0, -2.0, 0], [-4.0, 5.0, 0], [1.0, 1.0, 2.0]]
    origin = -0.5 * (n - 1) * vox
    brute = []
    half = 0.5 * n * vox * np.sqrt(3) + 2 * vox
    ss = np.arange(-half, half, 0.1)
    for M in mats:
        src = M[:3, :3] @ lat["src0"] + M[:3, 3]
        trg = np.einsum("ij,hwj->hwi", M[:3, :3], lat["trg0"]) + M[:3, 3]
        u = (trg - src) / np.linalg.norm(trg - src, axis=-1, keepdims=True)
        s0 = np.einsum("hwi,i->hw", u, -src)  # closest approach to the isocentre
        pts = src + (s0[..., None] + ss)[..., None] * u[:, :, None, :]
        val = map_coordinates(vol, ((pts - origin) / vox).reshape(-1, 3).T, order=1, mode="constant", cval=0.0)
        brute.append(val.reshape(pts.shape[:3]).sum(-1) * 0.1)
    brute = np.array(brute)
    yy, xx = np.mgrid[:H, :H]

    def centroid(a):
        return np.array([(a * yy).sum() / a.sum(), (a * xx).sum() / a.sum()])

    for br in (31, 24, 48):
        m = C.from_volume(vol, vox, (2, 2, 2), br)
        out = C.forward_pinhole(m, mats, lat)[:, 0]
        for k in range(3):
            assert np.abs(centroid(out[k]) - centroid(brute[k])).max() <= 0.3, (br, k)
            assert np.abs(out[k] - brute[k]).max() <= 0.12 * np.abs(brute[k]).max(), (br, k)
    m = C.from_volume(vol, vox, (2, 2, 2), 31)
    moved = dict(m, centers=m["centers"] + np.array([1.0, 0.0, 0.0]))
    out = C.forward_pinhole(moved, mats, lat)[:, 0]
    assert max(np.abs(centroid(out[k]) - centroid(brute[k])).max() for k in range(3)) > 0.3


# ---------------------------------------------------------------------------
# edge-dominant gradient NCC
# ---------------------------------------------------------------------------
def test_edge_stencils_equal_library_routines_and_adjoints():
    """Sobel x / y and the 3 x 3 box as scipy.ndimage.correlate (zero padding) and
    torch's conv2d / avg_pool2d (padding 1, count_include_pad) in float64; the adjoints
    by <K p, g> = <p, K^T g>."""
    from scipy.ndimage import correlate
    rng = np.random.default_rng(3)
    img = rng.standard_normal((2, 1, 13, 17))
    for k in (C.SOBEL_X, C.SOBEL_Y, C.BOX3):
        ref = np.stack([correlate(img[b, 0], k, mode="constant", cval=0.0) for b in range(2)])[:, None]
        np.testing.assert_allclose(C.correlate3(img, k), ref, atol=1e-13)
        g = rng.standard_normal(img.shape)
        assert abs(np.sum(C.correlate3(img, k) * g) - np.sum(img * C.correlate3_adjoint(g, k))) < 1e-10
    t = torch.from_numpy(img)
    kx = torch.tensor(C.SOBEL_X).view(1, 1, 3, 3)
    np.testing.assert_allclose(C.correlate3(img, C.SOBEL_X), torch.nn.functional.conv2d(t, kx, padding=1).numpy(),
                               atol=1e-13)
    np.testing.assert_allclose(C.correlate3(img, C.BOX3), torch.nn.functional.avg_pool2d(t, 3, 1, 1).numpy(),
                               atol=1e-13)


def _vignette_case(H=96):
    """The reference's vignetting case (tests/test_cone_beam_fvr.py:95-107), float64."""
    yy, xx = np.meshgrid(np.linspace(-1, 1, H), np.linspace(-1, 1, H), indexing="ij")
    img = np.exp(-((xx - 0.1) ** 2 + (yy + 0.05) ** 2) / 0.04) + 0.5 * ((xx ** 2 + yy ** 2) < 0.25)
    ref = img[None, None]
    vig = ref + 0.8 * xx[None, None]
    vig[:, :, :6, :] = -10.0
    vig[:, :, -6:, :] = -10.0
    shifted = np.roll(ref, (4, 4), axis=(-2, -1))
    return ref, vig, shifted


def test_edge_ncc_reference_known_answers_and_invariants():
    ref, vig, shifted = _vignette_case()
    assert C.edge_dominant_ncc(ref, ref) > 0.999
    assert C.edge_dominant_ncc(ref, vig) > 0.98
    assert C.edge_dominant_ncc(shifted, vig) < 0.50
    rng = np.random.default_rng(5)
    p = rng.standard_normal((3, 1, 100, 104)).cumsum(-1)  # rough images with edges
    t = rng.standard_normal((3, 1, 100, 104)).cumsum(-2)
    both = C.edge_dominant_ncc(p, t, reduction="none")
    assert both.shape == (3,)
    assert abs(C.edge_dominant_ncc(p, p) - 1.0) < 1e-12
    np.testing.assert_allclose(C.edge_dominant_ncc(p, 2.5 * t + 7.0, reduction="none"), both, atol=1e-12)  # affine
    np.testing.assert_allclose(C.edge_dominant_ncc(t, p, reduction="none"), both, atol=1e-12)  # symmetric
    np.testing.assert_allclose(C.edge_dominant_ncc(p, -t, reduction="none"), -both, atol=1e-12)
    assert abs(C.edge_dominant_ncc(p, t) - both.mean()) < 1e-15
    # the blurred term: (1 - w) fine + w coarse for W >= 96, the fine term alone below
    fine = C.edge_dominant_ncc(p, t, multiscale_weight=0.0, reduction="none")
    w1 = C.edge_dominant_ncc(p, t, multiscale_weight=1.0, reduction="none")
    np.testing.assert_allclose(both, 0.75 * fine + 0.25 * w1, atol=1e-12)
    assert not np.allclose(fine, w1)
    q, s = p[..., :95], t[..., :95]
    np.testing.assert_allclose(C.edge_dominant_ncc(q, s, reduction="none"),
                               C.edge_dominant_ncc(q, s, multiscale_weight=0.0, reduction="none"), atol=0)
    # the crop: rows / columns within crop of the border do not count (W // 8 = 13 here)
    p2 = p.copy()
    p2[..., :11, :] = rng.standard_normal(p2[..., :11, :].shape) * 50  # 11 = 13 - 2 (the blur + Sobel reach)
    np.testing.assert_allclose(C.edge_dominant_ncc(p2, t, reduction="none"), both, atol=1e-12)


@pytest.mark.parametrize("W,crop", [(40, None), (104, None), (100, 0)])
def test_edge_ncc_gradient_matches_directional_differences(W, crop):
    """The oracle's exact gradient w.r.t. the predicted images against FP64 central
    differences along random directions (both reductions; crop 0 reaches the zero
    padding at the border; W = 104 includes the blurred term)."""
    rng = np.random.default_rng(W)
    p = rng.standard_normal((2, 1, 36, W)).cumsum(-1)
    t = rng.standard_normal((2, 1, 36, W)).cumsum(-2)
    for red in ("none", "mean"):
        val, g = C.edge_dominant_ncc(p, t, border_crop=crop, reduction=red, with_grad=True)
        for _ in range(3):
            d = rng.standard_normal(p.shape)
            h = 1e-5
            fp = np.sum(C.edge_dominant_ncc(p + h * d, t, border_crop=crop, reduction=red))
            fm = np.sum(C.edge_dominant_ncc(p - h * d, t, border_crop=crop, reduction=red))
            fd = (fp - fm) / (2 * h)
            assert abs(fd - np.sum(g * d)) <= 1e-7 * max(1.0, abs(fd)), (red, fd, np.sum(g * d))
