"""GPU parity of the convention-N path (HIP kernels) against the FP64 oracle.

Tolerances (complex64 spectrum, FP32 interpolation and sums, float64 final
reductions): slices max|diff| <= 2e-6 * max|ref|; loss within 1e-5 relative;
gradients within 1e-4 of max|grad| (the target bound is 1e-3 relative).
"""

import numpy as np
import pytest
import torch

from oracle import diff_fvr_np as O

pytestmark = pytest.mark.gpu


def _vol(shape, seed=5):
    from fvr.phantom import ellipsoid_phantom
    n = max(shape)
    v = ellipsoid_phantom(n, n_ellipsoids=6, seed=seed).astype(np.float64)
    return np.ascontiguousarray(v[: shape[0], : shape[1], : shape[2]])


@pytest.fixture(scope="module", params=[(24, 24, 24), (17, 20, 23)], ids=["cube24", "ragged17x20x23"])
def setup(request):
    from fvr import DifferentiableFourierVolumeRenderer
    vol = _vol(request.param)
    r = DifferentiableFourierVolumeRenderer(vol)
    f3 = O.make_spectrum(vol)
    return vol, r, f3


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_spectrum_matches_oracle(setup):
    _, r, f3 = setup
    assert (r.pz, r.py, r.px) == f3.shape
    assert _rel(r.f_3d, f3) < 1e-12


@pytest.mark.parametrize("view", ["ap", "lateral", "axial"])
@pytest.mark.parametrize("pose", [(0, 0, 0, 0, 0), (7.5, -4.0, 12.0, 2.3, -1.7), (-15, 15, -15, -20, 20),
                                  (45, 30, -60, 0.5, 0.25)])
def test_slice_parity(setup, view, pose):
    _, r, f3 = setup
    got = r.extract_frequency_slice(*pose[:3], tx=pose[3], ty=pose[4], view=view)
    ref = O.extract_frequency_slice(f3, *pose[:3], tx=pose[3], ty=pose[4], view=view)
    assert got.shape == ref.shape
    assert _rel(got, ref) < 2e-6


def test_known_answer_mm_shift_is_np_roll():
    # the reference's known answer (tests/test_physical_units.py:28-34) on the GPU path
    from fvr import RobustDiffFVREngine
    z, y, x = np.indices((24,) * 3)
    vol = (np.exp(-((z - 10) ** 2 + (y - 11) ** 2 + (x - 10) ** 2) / 8)
           + .7 * np.exp(-((z - 15) ** 2 + (y - 15) ** 2 + (x - 15) ** 2) / 4))
    eng = RobustDiffFVREngine(vol, spacing=(2.0,) * 3)
    for view in ("ap", "lateral"):
        img = eng.render_calibrated_drr(view=view, crop_to_roi=False)
        shifted = eng.render_calibrated_drr(tx=4, ty=-2, view=view, crop_to_roi=False)
        # FP32 slice: 1e-6 of the image peak instead of the FP64 reference's 1e-9 absolute
        np.testing.assert_allclose(shifted, np.roll(img, (-1, 2), axis=(0, 1)), atol=1e-6 * np.abs(img).max())


@pytest.mark.parametrize("view", ["ap", "lateral", "axial"])
def test_render_drr_matches_oracle(view):
    """render_drr (diff_fvr.py:152-182: slice, centred inverse 2D FFT, scale, ReLU, centre
    crop) on the device against the oracle's FP64 render, with and without the crop, at an
    off-lattice pose with a sub-pixel translation: within 2e-5 of the image peak (FP32
    slice, whose samples are within 2e-6 of the largest; the inverse FFT runs in float64)."""
    from fvr import DifferentiableFourierVolumeRenderer
    vol = _vol((20, 24, 22))
    r = DifferentiableFourierVolumeRenderer(vol)
    f3 = O.make_spectrum(vol)
    for pose, crop in (((7.5, -4.0, 12.0, 2.3, -1.7), True), ((-3.0, 9.0, -6.5, -0.6, 3.1), False)):
        got = r.render_drr(*pose[:3], tx=pose[3], ty=pose[4], view=view, crop_to_roi=crop)
        ref = O.render_drr(f3, vol.shape, *pose[:3], tx=pose[3], ty=pose[4], view=view, crop_to_roi=crop)
        assert got.shape == ref.shape
        assert np.all(got >= 0.0)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= 2e-5, (view, pose, crop, err)


def _targets(f3, tp):
    T_ap = O.extract_frequency_slice(f3, *tp[:3], tx=tp[3], ty=tp[5], view="ap")
    T_la = O.extract_frequency_slice(f3, *tp[:3], tx=tp[4], ty=tp[5], view="lateral")
    return T_ap, T_la


POSES = [[3.3, 1.1, -2.2, 2.1, 1.5, -1.0], [-6.0, 4.0, 2.0, -3.0, 0.5, 2.5], [0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
         [14.0, -13.0, 9.0, 18.0, -19.0, 7.0]]


@pytest.mark.parametrize("pose", POSES)
def test_reference_fd_scheme_parity(setup, pose):
    from fvr import SpectralGradientEngine
    _, r, f3 = setup
    T_ap, T_la = _targets(f3, [1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
    eng = SpectralGradientEngine(r)
    L, g = eng.compute_biplanar_loss_and_gradients(np.array(pose), T_ap, T_la)
    Lr, gr = O.biplanar_loss_and_gradients(f3, np.array(pose), T_ap, T_la)
    assert abs(L - Lr) <= 1e-5 * max(abs(Lr), 1e-3)
    # rotation entries are FD quotients (NCC difference / 0.4 deg): an FP32 NCC error of
    # up to 1e-7 becomes 2.5e-7 in the gradient, independent of the gradient's size
    assert np.max(np.abs(g - gr)) <= 1e-4 * np.max(np.abs(gr)) + 2.5e-7, (g, gr)


@pytest.mark.parametrize("pose", POSES)
def test_analytic_adjoint_parity(setup, pose):
    from fvr import SpectralGradientEngine
    _, r, f3 = setup
    T_ap, T_la = _targets(f3, [1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
    eng = SpectralGradientEngine(r)
    L, g = eng.compute_biplanar_loss_and_gradients(np.array(pose), T_ap, T_la, rotation_grad="analytic")
    # the identity pose puts samples on lattice planes, where the interpolant has a kink: the
    # Hermitian-fold kernel's one-sided derivatives there are the oracle's fold convention
    fold = eng._evaluator(T_ap, T_la, True).herm
    Lr, gr = O.biplanar_loss_and_grad_analytic(f3, np.array(pose), T_ap, T_la, fold=fold)
    assert abs(L - Lr) <= 1e-5 * max(abs(Lr), 1e-3)
    assert np.max(np.abs(g - gr)) <= 1e-4 * np.max(np.abs(gr)) + 1e-9, (g, gr)


def test_batched_equals_single_and_is_deterministic(setup):
    from fvr import SpectralGradientEngine
    from fvr.phantom import random_poses
    _, r, f3 = setup
    T_ap, T_la = _targets(f3, [1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
    eng = SpectralGradientEngine(r)
    poses = random_poses(9, 15.0, 5.0, seed=3456)
    P = torch.from_numpy(poses).to(r.device)
    out_dev = eng.compute_biplanar_loss_and_gradients_batch(P, T_ap, T_la)
    assert all(isinstance(t, torch.Tensor) and t.is_cuda for t in out_dev)  # device in -> device out
    L1, g1 = [t.cpu().numpy().copy() for t in out_dev]
    L2, g2 = [t.copy() for t in eng.compute_biplanar_loss_and_gradients_batch(poses, T_ap, T_la)]
    np.testing.assert_array_equal(L1, L2)
    np.testing.assert_array_equal(g1, g2)
    for k in range(len(poses)):
        Ls, gs = eng.compute_biplanar_loss_and_gradients(poses[k], T_ap, T_la, rotation_grad="analytic")
        assert Ls == L1[k]
        np.testing.assert_array_equal(gs, g1[k])
        Lr, gr = O.biplanar_loss_and_grad_analytic(f3, poses[k], T_ap, T_la,
                                                   fold=eng._evaluator(T_ap, T_la, True).herm)
        assert abs(L1[k] - Lr) <= 1e-5 * max(abs(Lr), 1e-3)
        assert np.max(np.abs(g1[k] - gr)) <= 1e-4 * np.max(np.abs(gr)) + 1e-9


def test_zero_target_and_zero_volume_give_reference_zero_convention():
    from fvr import DifferentiableFourierVolumeRenderer, SpectralGradientEngine
    vol = _vol((16, 16, 16))
    r = DifferentiableFourierVolumeRenderer(vol)
    eng = SpectralGradientEngine(r)
    z = np.zeros((32, 32), dtype=np.complex128)
    L, g = eng.compute_biplanar_loss_and_gradients(np.ones(6), z, z)
    assert L == 1.0 and np.all(g == 0.0)
    L, g = eng.compute_biplanar_loss_and_gradients(np.ones(6), z, z, rotation_grad="analytic")
    assert L == 1.0 and np.all(g == 0.0)
    r0 = DifferentiableFourierVolumeRenderer(np.zeros((8, 8, 8)))
    s = r0.extract_frequency_slice(5, 5, 5, view="ap")
    assert np.all(s == 0)
    assert r0.compute_spectral_ncc(s, np.ones_like(s)) == 0.0


def test_pair_moment_helpers_match_oracle(setup):
    from fvr import SpectralGradientEngine
    _, r, f3 = setup
    a = O.extract_frequency_slice(f3, 2.0, 1.0, -1.0, tx=1.2, ty=-0.3, view="ap")
    b = O.extract_frequency_slice(f3, 1.0, 0.0, 0.5, tx=0.2, ty=0.3, view="ap")
    assert abs(r.compute_spectral_ncc(a, b) - O.compute_spectral_ncc(a, b)) < 1e-6
    got = SpectralGradientEngine(r).compute_spectral_ncc_and_grad_translation(a, b, view="ap")
    ref = O.ncc_and_grad_translation(a, b)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("storage,tol_slice,tol_grad", [("fp16", 2e-3, 2e-2), ("bf16", 2e-2, 1e-1)])
def test_reduced_precision_storage(storage, tol_slice, tol_grad):
    """FP16 / BF16 spectrum storage with one power-of-two scale, FP32 interpolation.

    Bounds (relative to the largest magnitude): slices 2e-3 (fp16, 11-bit significand)
    / 2e-2 (bf16, 8-bit); loss 1e-3 / 1e-2 absolute; gradient 2e-2 / 1e-1 of max|grad|.
    """
    from fvr import DifferentiableFourierVolumeRenderer, SpectralGradientEngine
    vol = _vol((24, 24, 24))
    r = DifferentiableFourierVolumeRenderer(vol, storage=storage)
    f3 = O.make_spectrum(vol)
    for view in ("ap", "lateral"):
        got = r.extract_frequency_slice(7.5, -4.0, 12.0, tx=2.3, ty=-1.7, view=view)
        ref = O.extract_frequency_slice(f3, 7.5, -4.0, 12.0, tx=2.3, ty=-1.7, view=view)
        assert _rel(got, ref) < tol_slice
    T_ap, T_la = _targets(f3, [1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
    eng = SpectralGradientEngine(r)
    for pose in POSES[:2]:
        L, g = eng.compute_biplanar_loss_and_gradients(np.array(pose), T_ap, T_la, rotation_grad="analytic")
        Lr, gr = O.biplanar_loss_and_grad_analytic(f3, np.array(pose), T_ap, T_la)
        assert abs(L - Lr) <= (1e-3 if storage == "fp16" else 1e-2)
        assert np.max(np.abs(g - gr)) <= tol_grad * np.max(np.abs(gr))


def test_device_setup_matches_host_setup(setup):
    """The device pose builder reproduces the host builder's items and Jacobians, and the
    device reduction + combine reproduces the host float64 combine."""
    from fvr import _lib
    from fvr.phantom import random_poses
    from fvr.sncc import BiplanarEvaluator
    _, r, f3 = setup
    T_ap, T_la = _targets(f3, [1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
    dev = r.device
    ev = BiplanarEvaluator(r.packed, torch.from_numpy(T_ap.astype(np.complex64)).to(dev),
                           torch.from_numpy(T_la.astype(np.complex64)).to(dev))
    poses = random_poses(6, 15.0, 5.0, seed=77)
    P = torch.from_numpy(poses).to(dev)
    ev.build_device(P)  # set-up only: the items are checked before any gather reads them
    got = ev.device_items(len(poses))
    ref, jac_ref, _ = ev.build(poses)
    for f in _lib.ITEM_DTYPE.names:
        if f in ("c", "ustep", "vstep"):
            # cos/sin of the two libms may differ in the last ulp: a few units of 2^-32 voxel
            assert np.max(np.abs(got[f].astype(np.int64) - ref[f].astype(np.int64))) <= 4, f
        elif f.startswith("ph_"):
            d = (got[f].astype(np.int64) - ref[f].astype(np.int64)) % (1 << 32)
            assert np.all((d <= 4) | (d >= (1 << 32) - 4)), f
        else:
            np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    np.testing.assert_allclose(ev.device_jacobians(len(poses)), jac_ref, atol=1e-15)
    L_d, g_d = [t.cpu().numpy() for t in ev.evaluate_device(P)]
    L_h, g_h = ev.evaluate_host(poses)
    np.testing.assert_allclose(L_d, L_h, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g_d, g_h, rtol=1e-6, atol=1e-9 * np.abs(g_h).max())


@pytest.mark.parametrize("storage", ["f32", "fp16"])
def test_hermitian_fold_equals_full_slice_kernel(setup, storage):
    """The Hermitian-fold kernel (half the samples) against the full-slice kernel on the
    same spectrum and targets: same sums up to float32 summation order."""
    from fvr.phantom import random_poses
    from fvr.sncc import BiplanarEvaluator
    _, r, f3 = setup
    if r.packed.dtype != storage:
        r.repack(storage)
    try:
        T_ap, T_la = _targets(f3, [1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
        dev = r.device
        ta = torch.from_numpy(T_ap.astype(np.complex64)).to(dev)
        tl = torch.from_numpy(T_la.astype(np.complex64)).to(dev)
        full = BiplanarEvaluator(r.packed, ta, tl, herm=False)
        fold = BiplanarEvaluator(r.packed, ta, tl, herm=True)
        assert fold.herm and not full.herm
        # generic poses: off the lattice planes both kernels differentiate the same function
        poses = random_poses(8, 15.0, 5.0, seed=91)
        L1, g1 = full.evaluate(poses)
        L2, g2 = fold.evaluate(poses)
        np.testing.assert_allclose(L2, L1, rtol=2e-6, atol=1e-7)
        assert np.max(np.abs(g2 - g1)) <= 2e-5 * np.abs(g1).max(), np.abs(g2 - g1).max()
        # the identity pose: equal losses (values are not affected by the kink)
        np.testing.assert_allclose(fold.evaluate(np.zeros((1, 6)))[0], full.evaluate(np.zeros((1, 6)))[0],
                                   rtol=2e-6, atol=1e-7)
    finally:
        if r.packed.dtype != "f32":
            r.repack("f32")


def _bf16_rne(x32: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bits, round to nearest even (finite inputs)."""
    u = x32.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_scaled_storage_equals_per_brick_quantisation(storage):
    """One scale per spectrum stores the numbers that a power-of-two scale per 8^3 brick
    stores (the FP16 / BF16 precision variant of the benchmark configurations): the
    device's own complex64 spectrum, quantised on the host per 8^3 brick of the storage
    lattice (max|component| / scale in [2^14, 2^15), round to nearest even), decodes to
    the same value, bit for bit, wherever the one-scale code is a normal number; the rest
    (fp16 subnormals under the one scale) differ by at most the subnormal step."""
    from fvr import DifferentiableFourierVolumeRenderer
    vol = _vol((20, 22, 18))
    r = DifferentiableFourierVolumeRenderer(vol, storage="f32")
    name = "zyx"
    d = r.packed.descs[r.packed.index[name]]
    sd = tuple(int(v) for v in d["sdim"])
    x = r.packed.storage[name].cpu().numpy().view(np.float32).reshape(sd + (2,))  # the packed complex64 source
    r.repack(storage)
    raw = r.packed.storage[name].cpu().numpy().view(np.uint16).reshape(sd + (2,))
    sc = float(r.packed.scales[name].cpu().numpy()[0])
    if storage == "fp16":
        got = raw.view(np.float16).astype(np.float64) * sc
    else:
        got = (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64) * sc
    # per-brick power-of-two scales over the storage lattice
    B = [(n + 7) // 8 for n in sd]
    pad = np.zeros((8 * B[0], 8 * B[1], 8 * B[2]), np.float64)
    pad[:sd[0], :sd[1], :sd[2]] = np.abs(x).max(axis=-1)
    bmax = pad.reshape(B[0], 8, B[1], 8, B[2], 8).max(axis=(1, 3, 5))
    e = np.where(bmax > 0, np.floor(np.log2(np.where(bmax > 0, bmax, 1.0))) - 14, 0)
    s_b = np.exp2(e)[np.ix_(*[np.arange(n) // 8 for n in sd])][..., None].astype(np.float32)
    q = (x / s_b).astype(np.float32)  # exact: power-of-two divisor, no underflow at these magnitudes
    if storage == "fp16":
        ref = q.astype(np.float16).astype(np.float64) * s_b
        normal = np.abs(x) / sc >= 2.0 ** -14
    else:
        ref = (_bf16_rne(q).astype(np.uint32) << 16).view(np.float32).astype(np.float64) * s_b
        normal = np.ones(x.shape, bool)
    assert normal.mean() > 0.5
    assert np.array_equal(got[normal], ref[normal])
    assert np.all(np.abs(got - ref)[~normal] <= 2.0 ** -24 * sc)


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_scaled_storage_image(storage):
    """The packed FP16/BF16 storage, read back and decoded by the header's definition:
    storage index s of an axis holds logical index s - 1 (zero outside the spectrum; with
    the wrap flag, logical index n holds index 0), quantised with the spectrum's one
    scale -- a power of two with max|component| / scale in [2^14, 2^15)."""
    from fvr import DifferentiableFourierVolumeRenderer, _lib
    vol = _vol((20, 22, 18))
    r = DifferentiableFourierVolumeRenderer(vol, storage=storage)
    f3 = O.make_spectrum(vol)
    cmax = max(np.abs(f3.real).max(), np.abs(f3.imag).max())
    for name, perm in (("zyx", (0, 1, 2)), ("zxy", (0, 2, 1))):
        k = r.packed.index[name]
        d = r.packed.descs[k]
        n, sd = np.array(d["n"]), np.array(d["sdim"])
        assert d["row_bytes"] == sd[2] * 4 and d["plane_bytes"] == sd[1] * sd[2] * 4
        raw = r.packed.storage[name].cpu().numpy().view(np.uint16).reshape(sd[0], sd[1], sd[2], 2)
        sc = r.packed.scales[name].cpu().numpy()
        assert sc.shape == (1,)
        sc = float(sc[0])
        spec = np.transpose(f3, perm)
        lg = [np.arange(m) - 1 for m in sd]  # logical index of each storage index
        if int(d["flags"]) & _lib.LAYOUT_WRAP:
            lg = [np.where(x == m, 0, x) for x, m in zip(lg, n)]
        ok = [(x >= 0) & (x < m) for x, m in zip(lg, n)]
        idx = np.ix_(*[np.clip(x, 0, m - 1) for x, m in zip(lg, n)])
        exp = spec[idx] * (ok[0][:, None, None] & ok[1][None, :, None] & ok[2][None, None, :])
        if storage == "fp16":
            dec = raw.view(np.float16).astype(np.float64)
        else:
            dec = (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        got = (dec[..., 0] + 1j * dec[..., 1]) * sc
        bits = 11 if storage == "fp16" else 8
        err = np.abs(got - exp)
        # round to nearest: half an ulp of the significand, plus fp16's subnormal step (2^-24 x scale)
        assert np.all(err <= 2.0 ** -bits * np.abs(exp) * (1 + 1e-12) + 2.0 ** -24 * sc)
        # the scale: a power of two with the spectrum's maximum in [2^14, 2^15) x scale
        assert np.log2(sc) == np.round(np.log2(sc))
        assert 2.0 ** 14 * (1 - 1e-6) <= cmax / sc < 2.0 ** 15


def test_robust_engine_registration_clean_and_metal():
    """register_biplanar_robust as the reference's integration test exercises it
    (tests/test_core_engine.py:13-73), on a seeded 64^3 ellipsoid phantom at 1.5 mm instead
    of its DICOM series: DRR shapes, the saturation mask, Powell from (1.5, -1, 0.8 deg,
    2, -1.5, 1.8 mm) to the identity pose on clean DRR targets (mTRE < 1 mm; each
    evaluation one fused loss-only launch) and on targets with a saturated patch (metal
    detected, masked spatial NCC, mTRE < 2 mm)."""
    from fvr import RobustDiffFVREngine
    from fvr.phantom import ellipsoid_phantom
    from fvr.registration_3d2d import compute_mtre_6dof
    n, vox = 64, 1.5
    vol = ellipsoid_phantom(n, n_ellipsoids=12, seed=11).astype(np.float64)
    eng = RobustDiffFVREngine(vol, spacing=(vox,) * 3, pad_factor=2.0)
    drr_ap = eng.render_calibrated_drr(0, 0, 0, tx=0, ty=0, view="ap")
    drr_lat = eng.render_calibrated_drr(0, 0, 0, tx=0, ty=0, view="lateral")
    assert drr_ap.shape == (n, n) and drr_lat.shape == (n, n)
    drr_metal = np.copy(drr_lat)
    drr_metal[10:18, 10:25] = np.max(drr_lat) * 5.0
    mask = eng.detect_metal_hardware_mask(drr_metal, threshold=0.98)
    assert mask.any() and not mask.all()
    true_pose, init = (0.0,) * 6, (1.5, -1.0, 0.8, 2.0, -1.5, 1.8)
    res = eng.register_biplanar_robust(drr_ap, drr_lat, initial_pose=init, max_iter=30)
    m_clean = compute_mtre_6dof(true_pose, tuple(res["final_pose"]), vol.shape, spacing=(vox,) * 3,
                                translation_units="mm")
    assert not res["metal_detected"] and m_clean < 1.0, (m_clean, res)
    ap_occ, lat_occ = np.copy(drr_ap), np.copy(drr_lat)
    ap_occ[5:20, 10:30] = np.max(drr_ap) * 5.0
    lat_occ[5:20, 10:30] = np.max(drr_lat) * 5.0
    res_m = eng.register_biplanar_robust(ap_occ, lat_occ, initial_pose=init, handle_metal=True, max_iter=30)
    m_metal = compute_mtre_6dof(true_pose, tuple(res_m["final_pose"]), vol.shape, spacing=(vox,) * 3,
                                translation_units="mm")
    assert res_m["metal_detected"] and m_metal < 2.0, (m_metal, res_m)
    print(f"clean: {res['evaluations']} evals {res['time_sec']:.2f} s mTRE {m_clean:.4f} mm; "
          f"metal: {res_m['evaluations']} evals {res_m['time_sec']:.2f} s mTRE {m_metal:.4f} mm")


def test_reference_numerical_gradient_case():
    """The reference's numerical-gradient case (tests/test_gradient_numerical.py:12-57): a
    32^3 cube holding a 12^3 unit box, targets at the identity pose, the biplanar loss and
    gradient at (1, -1, 0.5 deg, 1.5, -1, 2 voxels) against central differences (eps 1e-4)
    of the same loss.  The differences are taken in float64 with the oracle (the device's
    float32 loss is too coarse for eps 1e-4): the exact gradient (rotation_grad="analytic")
    against them, and the default gradient (the reference's scheme: closed-form translations,
    rotations differenced at +-0.2 deg) against the oracle's restatement of that scheme."""
    from fvr import DifferentiableFourierVolumeRenderer, SpectralGradientEngine
    cube = np.zeros((32, 32, 32), dtype=np.float64)
    cube[10:22, 10:22, 10:22] = 1.0
    r = DifferentiableFourierVolumeRenderer(cube, pad_factor=2.0)
    eng = SpectralGradientEngine(r)
    f3 = O.make_spectrum(cube)
    T_ap = O.extract_frequency_slice(f3, 0, 0, 0, tx=0, ty=0, view="ap")
    T_la = O.extract_frequency_slice(f3, 0, 0, 0, tx=0, ty=0, view="lateral")
    pose = np.array([1.0, -1.0, 0.5, 1.5, -1.0, 2.0])

    def loss64(p):
        yaw, pitch, roll, tx, ty, tz = p
        a = O.extract_frequency_slice(f3, yaw, pitch, roll, tx=tx, ty=tz, view="ap")
        b = O.extract_frequency_slice(f3, yaw, pitch, roll, tx=ty, ty=tz, view="lateral")
        return 0.5 * (1.0 - O.compute_spectral_ncc(a, T_ap)) + 0.5 * (1.0 - O.compute_spectral_ncc(b, T_la))

    eps = 1e-4
    g_fd = np.array([(loss64(pose + eps * e) - loss64(pose - eps * e)) / (2 * eps) for e in np.eye(6)])
    L, g = eng.compute_biplanar_loss_and_gradients(pose, T_ap, T_la, rotation_grad="analytic")
    assert abs(L - loss64(pose)) <= 1e-6
    err = np.abs(g - g_fd).max() / np.abs(g_fd).max()
    assert err <= 2e-3, (g, g_fd, err)
    L, g = eng.compute_biplanar_loss_and_gradients(pose, T_ap, T_la)
    Lr, gr = O.biplanar_loss_and_gradients(f3, pose, T_ap, T_la, delta_rot_deg=0.2)
    assert abs(L - Lr) <= 1e-6
    assert np.abs(g - gr).max() <= 1e-4 * np.abs(gr).max(), (g, gr)


def test_register_adam_follows_the_oracle():
    """GradientRegistrationOptimizer.register_adam (gradient_optimizer.py:24-90) from the
    reference benchmark's start (tests/test_gradient_optimizer.py:25-47: 2.5, -3, 1.5 deg,
    3.5, -2, 4) on a seeded 40^3 phantom, targets at the identity: the first 5 steps follow
    the oracle's float64 Adam over its float64 loss and gradient (same hyper-parameters) to
    1e-3 in every pose entry, and 30 steps lower the loss and the mTRE."""
    from fvr import DifferentiableFourierVolumeRenderer
    from fvr.gradient_optimizer import GradientRegistrationOptimizer
    from fvr.registration_3d2d import compute_mtre_6dof
    vol = _vol((40, 40, 40), seed=9)
    r = DifferentiableFourierVolumeRenderer(vol)
    f3 = O.make_spectrum(vol)
    T_ap = O.extract_frequency_slice(f3, 0, 0, 0, tx=0, ty=0, view="ap")
    T_la = O.extract_frequency_slice(f3, 0, 0, 0, tx=0, ty=0, view="lateral")
    init = np.array([2.5, -3.0, 1.5, 3.5, -2.0, 4.0])
    opt = GradientRegistrationOptimizer(r)
    got = opt.register_adam(T_ap, T_la, init_pose=init, max_iter=5, tol=0.0)
    ref = O.adam_register(lambda p: O.biplanar_loss_and_gradients(f3, p, T_ap, T_la), init, max_iter=5, tol=0.0)
    for (s1, l1, p1), (s2, l2, p2) in zip(got["trajectory"], ref["trajectory"]):
        assert s1 == s2 and abs(l1 - l2) <= 1e-5 and np.abs(p1 - p2).max() <= 1e-3, (s1, p1, p2)
    res = opt.register_adam(T_ap, T_la, init_pose=init, max_iter=30)
    traj = res["trajectory"]
    assert res["final_loss"] < traj[0][1]
    m0 = compute_mtre_6dof((0.0,) * 6, tuple(init), vol.shape)
    m1 = compute_mtre_6dof((0.0,) * 6, tuple(res["final_pose"]), vol.shape)
    assert m1 < m0, (m0, m1)


def test_tall_volume_takes_the_full_slice_kernel():
    """A volume taller than it is wide (24 x 16 x 14: AP slices 48 x 28, lateral 48 x 32,
    h > w) cannot be folded: the evaluator picks the full-slice kernel by itself, and the
    loss and exact gradient match the FP64 oracle as on foldable shapes."""
    from fvr import DifferentiableFourierVolumeRenderer, SpectralGradientEngine
    vol = _vol((24, 16, 14), seed=6)
    r = DifferentiableFourierVolumeRenderer(vol)
    f3 = O.make_spectrum(vol)
    T_ap, T_la = _targets(f3, [1.0, -2.0, 1.5, 0.7, -1.2, 0.4])
    eng = SpectralGradientEngine(r)
    assert not eng._evaluator(T_ap, T_la, True).herm
    for pose in ([3.0, -2.0, 1.0, 1.5, -1.0, 0.5], [-7.5, 4.0, 9.0, -0.6, 2.2, -1.3]):
        L, g = eng.compute_biplanar_loss_and_gradients(np.array(pose), T_ap, T_la, rotation_grad="analytic")
        Lr, gr = O.biplanar_loss_and_grad_analytic(f3, np.array(pose), T_ap, T_la)
        assert abs(L - Lr) <= 1e-5 * max(abs(Lr), 1e-3)
        assert np.max(np.abs(g - gr)) <= 1e-4 * np.max(np.abs(gr)) + 1e-9, (g, gr)


def test_batched_powell_registration_matches_scipy_driver():
    """register_biplanar_robust(optimizer="batched_powell") on the clean objective: every
    line-search round is one fused loss-only launch over a grid of poses.  From the
    reference integration test's start (1.5, -1, 0.8 deg, 2, -1.5, 1.8 mm) on a seeded 64^3
    phantom at 1.5 mm it recovers the pose as the scipy driver does (mTRE < 1 mm, the
    reference's bound), with far fewer launches than evaluations."""
    from fvr import RobustDiffFVREngine
    from fvr.phantom import ellipsoid_phantom
    from fvr.registration_3d2d import compute_mtre_6dof
    n, vox = 64, 1.5
    vol = ellipsoid_phantom(n, n_ellipsoids=12, seed=11).astype(np.float64)
    eng = RobustDiffFVREngine(vol, spacing=(vox,) * 3, pad_factor=2.0)
    drr_ap = eng.render_calibrated_drr(0, 0, 0, tx=0, ty=0, view="ap")
    drr_lat = eng.render_calibrated_drr(0, 0, 0, tx=0, ty=0, view="lateral")
    init = (1.5, -1.0, 0.8, 2.0, -1.5, 1.8)
    out = {}
    for opt in ("powell", "batched_powell"):
        res = eng.register_biplanar_robust(drr_ap, drr_lat, initial_pose=init, max_iter=30, optimizer=opt)
        m = compute_mtre_6dof((0.0,) * 6, tuple(res["final_pose"]), vol.shape, spacing=(vox,) * 3,
                              translation_units="mm")
        out[opt] = (m, res)
        print(f"{opt}: {res['evaluations']} evals in {res['launches']} launches, {res['time_sec']:.3f} s, "
              f"{res['evaluations'] / res['time_sec']:.0f} evals/s, mTRE {m:.4f} mm, loss {res['final_loss']:.3e}")
        assert not res["metal_detected"] and m < 1.0, (opt, m, res)
    m_b, r_b = out["batched_powell"]
    assert r_b["launches"] * 4 < r_b["evaluations"]
    assert r_b["final_loss"] <= out["powell"][1]["final_loss"] * 1.5 + 1e-6
    with pytest.raises(ValueError):
        eng.register_biplanar_robust(drr_ap, drr_lat, optimizer="nelder-mead")
