"""GPU parity at the benchmark workloads' full sizes, against the FP64 oracle.

* 256^3 volume -> 512^3 spectrum, AP + lateral 512^2 slices, 64 candidate poses
  (rotations U(-15, 15) deg, translations U(-20, 20) voxels) in one launch of the
  Hermitian-fold kernel with the autotuned plan, FP16 and complex64 storage:
  loss and exact 6-DoF gradient of 8 candidates taken from inside the full launch
  against ``biplanar_loss_and_grad_analytic``; one candidate against FP64 central
  differences (h = 1e-3) of the oracle's loss; the device target preparation
  against the oracle's.
* PoU-FVR: 256^3 volume, 4 x 4 x 4 blocks with 25 % overlap resampled to 64^3
  (P = 128: the fused LDS slice + inverse FFT and the fused adjoint at their
  largest LDS footprint), pinhole SDD 1020 / SAD 510 mm, 256^2 detector at 0.8 mm,
  64 poses in one launch: slab images, projections and the NCC-loss pose gradient
  against the oracle and its FP64 central differences.
* Ragged multistart: 128^3 / 192^3 / 256^3 volumes (256^3 / 384^3 / 512^3 FP16
  spectra), 512 starting poses split 96 / 160 / 256 in one fused launch: poses of
  every volume against the oracle, the 100-step Adam loop replayed from a HIP graph
  against the eager loop, bit for bit, and the recovered poses.

Tolerances: complex64 storage -- loss within 1e-5 relative, gradient within 1e-4
of max|grad| (as at small sizes); FP16 storage (11-bit significand, FP32
interpolation and sums) -- loss within 1e-4 absolute, gradient within 1e-3 of
max|grad| per pose; finite differences 2e-3 of max|grad|.  PoU float32 block
spectra: slab images and projections within 1e-4 of max|ref|; BF16 block spectra
within 2e-2 of max|ref| (8-bit significand).
"""

import numpy as np
import pytest
import torch

from oracle import cone_beam_np as C
from oracle import diff_fvr_np as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

DEV = torch.device("cuda", 0)


def _c64(x) -> np.ndarray:
    """complex64 values (what the device holds), as complex128 for the oracle."""
    return np.asarray(x).astype(np.complex64).astype(np.complex128)


def _oracle_targets(f3, orig_shape, tp):
    """AP / lateral target spectra of DRRs rendered at pose tp (diff_fvr.py:152-182, core_engine.py:71-85)."""
    yaw, pitch, roll, tx, ty, tz = [float(v) for v in tp]
    out = []
    for view, a in (("ap", tx), ("lateral", ty)):
        img = O.render_drr(f3, orig_shape, yaw, pitch, roll, tx=a, ty=tz, view=view, crop_to_roi=True)
        h, w = O.slice_dims(f3.shape, view)
        out.append(O.prepare_spectrum(img, h, w))
    return out


# ---------------------------------------------------------------------------
# 256^3 volume, 512^3 spectrum, 64 candidates
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def headline():
    from fvr import DifferentiableFourierVolumeRenderer
    from fvr.phantom import ellipsoid_phantom, random_poses
    vol = ellipsoid_phantom(256, n_ellipsoids=12, seed=1234)
    rng = np.random.default_rng(2345)
    true_pose = np.concatenate([rng.uniform(-5, 5, 3), rng.uniform(-5, 5, 3)])
    cands = random_poses(64, 15.0, 20.0, seed=3456)
    print("[full-size] oracle spectrum 512^3", flush=True)
    f3 = O.make_spectrum(vol)
    T_ap, T_lat = _oracle_targets(f3, vol.shape, true_pose)
    r = DifferentiableFourierVolumeRenderer(vol, pad_factor=2.0, storage="f32", device=DEV)
    return dict(vol=vol, f3=f3, tp=true_pose, cands=cands, T=(T_ap, T_lat), r=r)


def test_256cube_spectrum_and_target_preparation_match_oracle(headline):
    """a1 and a7 at full size: the device's float64 spectrum (rocFFT) against numpy's,
    and prepare_spectrum_device (core_engine.py:71-85) of the oracle's 256^2 DRRs."""
    from fvr.core_engine import prepare_spectrum_device
    f3, r = headline["f3"], headline["r"]
    assert (r.pz, r.py, r.px) == f3.shape == (512, 512, 512)
    err = float((r._spec - torch.from_numpy(f3).to(DEV)).abs().max()) / float(np.abs(f3).max())
    assert err < 1e-12, err
    yaw, pitch, roll, tx, ty, tz = headline["tp"]
    for k, (view, a) in enumerate((("ap", tx), ("lateral", ty))):
        img = O.render_drr(f3, headline["vol"].shape, yaw, pitch, roll, tx=a, ty=tz, view=view)
        assert img.shape == (256, 256)
        got = prepare_spectrum_device(img, 512, 512, DEV).cpu().numpy()
        ref = headline["T"][k]
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        # the whole device chain (FP32 slice of the complex64 storage -> float64 inverse FFT -> DRR)
        drr = r.render_drr_device(yaw, pitch, roll, tx=a, ty=tz, view=view, crop_to_roi=True)
        dev_t = prepare_spectrum_device(drr, 512, 512, DEV).cpu().numpy()
        assert np.abs(dev_t - ref).max() <= 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("storage", ["fp16", "f32"])
def test_256cube_64cand_launch_matches_oracle(headline, storage):
    from fvr.multistart import autotune
    from fvr.sncc import BiplanarEvaluator
    r, f3, cands = headline["r"], headline["f3"], headline["cands"]
    if r.storage != storage:
        r.repack(storage)
    T_ap, T_lat = (torch.from_numpy(t.astype(np.complex64)).to(DEV) for t in headline["T"])
    ta, tl = (_c64(t) for t in headline["T"])
    ev = BiplanarEvaluator(r.packed, T_ap, T_lat, adjoint=True)
    assert ev.herm
    P = torch.from_numpy(np.ascontiguousarray(cands)).to(DEV)
    plan, _ = autotune(ev, P, reps=2)
    L, G = (x.cpu().numpy().copy() for x in ev.evaluate_device(P))
    assert np.all(np.isfinite(L)) and np.all(np.isfinite(G))
    errs = []
    for k in (0, 9, 17, 26, 38, 45, 51, 63):
        Lr, gr = O.biplanar_loss_and_grad_analytic(f3, cands[k], ta, tl, fold=True)
        gmax = np.abs(gr).max()
        if storage == "f32":
            assert abs(L[k] - Lr) <= 1e-5 * abs(Lr), (k, L[k], Lr)
            assert np.abs(G[k] - gr).max() <= 1e-4 * gmax, (k, G[k], gr)
        else:
            assert abs(L[k] - Lr) <= 1e-4, (k, L[k], Lr)
            assert np.abs(G[k] - gr).max() <= 1e-3 * gmax, (k, G[k], gr)
        errs.append((abs(L[k] - Lr), np.abs(G[k] - gr).max() / gmax))
    e = np.array(errs)
    print(f"[full-size] {storage} plan {plan}: max |dL| {e[:, 0].max():.2e}, max rel dgrad {e[:, 1].max():.2e}")
    # the plan is a layout choice only: the default plan gives the same sums to float32 regrouping
    from fvr.sncc import LaunchPlan
    ev.set_plan(LaunchPlan())
    L2, G2 = (x.cpu().numpy() for x in ev.evaluate_device(P))
    np.testing.assert_allclose(L2, L, rtol=1e-5)
    assert np.abs(G2 - G).max() <= 1e-4 * np.abs(G).max()


def test_256cube_gradient_matches_fp64_central_differences(headline):
    """The exact adjoint of the full launch against FP64 central differences (h = 1e-3 deg /
    voxel) of the oracle's loss at one candidate: 2e-3 of max|grad|."""
    from fvr.sncc import BiplanarEvaluator
    r, f3, cands = headline["r"], headline["f3"], headline["cands"]
    if r.storage != "f32":
        r.repack("f32")
    T_ap, T_lat = (torch.from_numpy(t.astype(np.complex64)).to(DEV) for t in headline["T"])
    ta, tl = (_c64(t) for t in headline["T"])
    ev = BiplanarEvaluator(r.packed, T_ap, T_lat, adjoint=True)
    L, G = ev.evaluate(cands)
    k = 5
    h = 1e-3
    g_fd = np.array([(O.biplanar_loss(f3, cands[k] + h * e, ta, tl) - O.biplanar_loss(f3, cands[k] - h * e, ta, tl))
                     / (2 * h) for e in np.eye(6)])
    err = np.abs(G[k] - g_fd).max() / np.abs(g_fd).max()
    assert err <= 2e-3, (G[k], g_fd, err)


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_512cube_one_scale_storage_equals_per_brick_quantisation(headline, storage):
    """The 512^3 spectrum stored with one power-of-two scale holds the numbers that a
    power-of-two scale per 8^3 brick of the storage lattice holds (max|component| / scale
    in [2^14, 2^15), round to nearest even, quantised on the device from the same
    complex64 values): bit for bit wherever the one-scale FP16 code is a normal number,
    and within FP16's subnormal step (2^-24 x scale) elsewhere.  BF16 has FP32's exponent
    range, so every element is equal.  Prints the subnormal fraction and its energy share."""
    r = headline["r"]
    if r.storage != "f32":
        r.repack("f32")
    name = "zyx"
    d = r.packed.descs[r.packed.index[name]]
    sd = tuple(int(v) for v in d["sdim"])
    x = r.packed.storage[name].view(torch.float32).reshape(sd + (2,)).clone()
    r.repack(storage)
    sc = float(r.packed.scales[name][0])
    raw = r.packed.storage[name].view(torch.float16 if storage == "fp16" else torch.bfloat16).reshape(sd + (2,))
    got = raw.double() * sc
    nb = [(n + 7) // 8 for n in sd]
    pad = torch.zeros((8 * nb[0], 8 * nb[1], 8 * nb[2]), dtype=torch.float32, device=DEV)
    pad[:sd[0], :sd[1], :sd[2]] = x.abs().amax(-1)
    bmax = pad.view(nb[0], 8, nb[1], 8, nb[2], 8).amax(dim=(1, 3, 5)).double()
    e = torch.where(bmax > 0, torch.floor(torch.log2(torch.where(bmax > 0, bmax, 1.0))) - 14, 0.0)
    del pad, bmax
    ix = [torch.arange(n, device=DEV) // 8 for n in sd]
    s_b = torch.exp2(e)[ix[0]][:, ix[1]][:, :, ix[2]].float()[..., None]
    q = x / s_b  # exact: a power-of-two divisor, no underflow at these magnitudes
    ref = (q.half() if storage == "fp16" else q.bfloat16()).double() * s_b
    del q
    normal = (x.abs() / sc >= 2.0 ** -14) if storage == "fp16" else torch.ones_like(x, dtype=torch.bool)
    frac = 1.0 - float(normal.double().mean())
    x2 = x.double() ** 2
    energy = float(x2[~normal].sum() / x2.sum())
    print(f"[full-size] 512^3 {storage}: scale 2^{int(np.log2(sc))}, subnormal fraction {frac:.3e}, "
          f"their energy share {energy:.3e}")
    assert torch.equal(got[normal], ref[normal])
    sub = (got - ref).abs()[~normal]
    assert sub.numel() == 0 or float(sub.max()) <= 2.0 ** -24 * sc


# ---------------------------------------------------------------------------
# PoU-FVR: 256^3, 4 x 4 x 4 blocks, P = 128, 256^2 pinhole detector, 64 poses
# ---------------------------------------------------------------------------
SDD, SAD, DET, PIX = 1020.0, 510.0, 256, 0.8


def _k_inv():
    f = SDD / PIX
    return np.array([[1.0 / f, 0.0, -DET / (2 * f)], [0.0, 1.0 / f, -DET / (2 * f)], [0.0, 0.0, 1.0]])


def _pou_mats(rot_deg, trans_mm):
    from scipy.spatial.transform import Rotation
    M = np.tile(np.eye(4), (len(rot_deg), 1, 1))
    M[:, :3, :3] = Rotation.from_euler("zyx", rot_deg, degrees=True).as_matrix()
    M[:, :3, 3] = np.array([0.0, 0.0, -SAD]) + trans_mm
    return M


@pytest.fixture(scope="module")
def pou():
    from fvr import PerspectiveSlabFVR
    from fvr.phantom import ellipsoid_phantom
    vol = ellipsoid_phantom(256, n_ellipsoids=12, seed=1234)
    print("[full-size] PoU oracle blocks (55 x 128^3)", flush=True)
    o = C.from_volume(vol.astype(np.float64), 1.0, preset="isotropic")
    m32 = PerspectiveSlabFVR.from_volume(vol, vox_mm=1.0, preset="isotropic", dtype=torch.float32, device=DEV)
    m16 = PerspectiveSlabFVR.from_volume(vol, vox_mm=1.0, preset="isotropic", device=DEV)
    lat = m32.make_pinhole_lattice(torch.from_numpy(_k_inv()).float().to(DEV), SDD, DET, DET)
    lat_np = C.make_pinhole_lattice(_k_inv(), SDD, DET, DET)
    rng = np.random.default_rng(3456)
    true = _pou_mats(rng.uniform(-5, 5, (1, 3)), rng.uniform(-5, 5, (1, 3)))
    cands = _pou_mats(rng.uniform(-5, 5, (64, 3)), rng.uniform(-5, 5, (64, 3)))
    return dict(o=o, m32=m32, m16=m16, lat=lat, lat_np=lat_np, true=true, cands=cands)


def test_pou_256cube_blocks_match_oracle(pou):
    o, m = pou["o"], pou["m32"]
    assert (m.B, m.P) == (o["B"], o["P"]) == (55, 128) and abs(m.ds_block - o["ds"]) < 1e-12
    assert m.fused_ok()
    np.testing.assert_allclose(m.centers.cpu().numpy(), o["centers"], atol=1e-4)
    F = m.F_blk_ri.cpu().numpy().astype(np.float64)
    ref = o["F_blk"]
    assert np.abs(F[:, 0] + 1j * F[:, 1] - ref).max() <= 1e-6 * np.abs(ref).max()


def test_pou_256cube_slab_images_and_projections_match_oracle(pou):
    o, lat, lat_np, cands = pou["o"], pou["lat"], pou["lat_np"], pou["cands"]
    sel = [3, 40]
    R = cands[sel, :3, :3]
    u = np.einsum("nij,j->ni", R, lat_np["u0"])
    v = np.einsum("nij,j->ni", R, lat_np["v0"])
    ref_imgs = C.slab_images(o, u, v)
    ref_out = C.forward_pinhole(o, cands[sel], lat_np)
    M = torch.from_numpy(cands).float().to(DEV)
    for name, tol in (("m32", 1e-4), ("m16", 2e-2)):
        m = pou[name]
        imgs = m._slice_and_ifft(torch.from_numpy(u).float(), torch.from_numpy(v).float()).cpu().numpy()
        assert imgs.shape == ref_imgs.shape == (55, 2, 128, 128)
        e_img = np.abs(imgs - ref_imgs).max() / np.abs(ref_imgs).max()
        with torch.no_grad():
            out = m.forward_pinhole(M, lat).cpu().numpy()  # all 64 poses in one launch
        assert out.shape == (64, 1, 256, 256) and np.all(np.isfinite(out))
        e_out = np.abs(out[sel] - ref_out).max() / np.abs(ref_out).max()
        print(f"[full-size] PoU {name}: slab images {e_img:.2e}, projections {e_out:.2e}")
        assert e_img <= tol and e_out <= tol, (name, e_img, e_out)


def test_pou_256cube_mirror_queue_and_per_thread_pass_agree(pou):
    """At P = 128 with 64 poses (bands of explicitly gathered mirrors along the box faces in
    every slice): the compacted mirror queue and the per-thread pass give the same slab images
    bit for bit and the same pose gradient up to float rounding."""
    from fvr import _lib
    m, lat, cands = pou["m32"], pou["lat"], pou["cands"]
    R = torch.from_numpy(cands[:, :3, :3]).float()
    lat0 = pou["lat_np"]
    u = torch.einsum("nij,j->ni", R, torch.from_numpy(lat0["u0"]).float())
    v = torch.einsum("nij,j->ni", R, torch.from_numpy(lat0["v0"]).float())
    res = []
    for on in (1, 0):
        prev = _lib.lib().fvr_pou_set_mirror_queue(on)
        try:
            imgs = m._slice_and_ifft(u, v).cpu().numpy()
            Mt = torch.from_numpy(cands).float().to(DEV).requires_grad_(True)
            m.forward_pinhole(Mt, lat).square().sum().backward()
            res.append((imgs, Mt.grad.double().cpu().numpy()))
        finally:
            _lib.lib().fvr_pou_set_mirror_queue(prev)
    assert np.array_equal(res[0][0], res[1][0])
    err = np.abs(res[0][1] - res[1][1]).max() / np.abs(res[1][1]).max()
    print(f"[full-size] PoU mirror queue vs per-thread pass: gradient {err:.2e}")
    assert err <= 1e-5


def test_pou_256cube_fused_path_matches_generic_composition(pou):
    """At P = 128 (the register transforms of the forward and the adjoint, the gather-form warp
    adjoint, the mirror queue): projections and the pose gradient of a weighted sum for 4 poses
    against the generic composition (slice operator, torch.fft, warp kernels, autograd), which
    shares none of the fused kernels' transforms; projections within 1e-5 and gradients within
    1e-4 of max (float32 spectra; measured 1.5e-6 and 3.8e-6)."""
    m, lat, cands = pou["m32"], pou["lat"], pou["cands"][:4]
    w = torch.randn((4, 1, DET, DET), generator=torch.Generator().manual_seed(21)).to(DEV)
    res = {}
    for fused in (True, False):
        m.fused_kernels = fused
        try:
            Mt = torch.from_numpy(cands).float().to(DEV).requires_grad_(True)
            out = m.forward_pinhole(Mt, lat)
            (out * w).sum().backward()
            res[fused] = (out.detach().double().cpu().numpy(), Mt.grad.double().cpu().numpy())
        finally:
            m.fused_kernels = True
    e_out = np.abs(res[True][0] - res[False][0]).max() / np.abs(res[False][0]).max()
    e_g = np.abs(res[True][1] - res[False][1]).max() / np.abs(res[False][1]).max()
    print(f"[full-size] PoU fused vs generic: projections {e_out:.2e}, pose gradient {e_g:.2e}")
    assert e_out <= 1e-5 and e_g <= 1e-4, (e_out, e_g)


def _ncc_loss_torch(out, t0c):
    o0 = out - out.mean(dim=(-2, -1), keepdim=True)
    ncc = (o0 * t0c).sum(dim=(-2, -1)) / (o0.norm(dim=(-2, -1)) * t0c.norm() + 1e-8)
    return (1.0 - ncc).sum()


def test_pou_256cube_pose_gradient_matches_oracle_fd(pou):
    """The fused adjoint at P = 128 (64 poses in one launch, NCC loss against a seeded
    target) against FP64 central differences of the oracle's loss at one pose, every
    entry of its 4 x 4 matrix (h = 1e-2 mm for translations, 1e-4 for rotation entries);
    the BF16 model's gradient against the float32 model's."""
    o, lat, lat_np, cands = pou["o"], pou["lat"], pou["lat_np"], pou["cands"]
    m = pou["m32"]
    with torch.no_grad():
        target = m.forward_pinhole(torch.from_numpy(pou["true"]).float().to(DEV), lat)
    t0c = target - target.mean()
    grads = {}
    for name in ("m32", "m16"):
        Mt = torch.from_numpy(cands).float().to(DEV).requires_grad_(True)
        _ncc_loss_torch(pou[name].forward_pinhole(Mt, lat), t0c).backward()
        grads[name] = Mt.grad.double().cpu().numpy()
        assert np.all(np.isfinite(grads[name]))
    tgt = target.double().cpu().numpy()[0, 0]
    k = 7

    def loss(Mk):
        return C.ncc_loss(C.forward_pinhole(o, Mk[None], lat_np)[0, 0], tgt)

    g = grads["m32"][k]
    fd = np.zeros((4, 4))
    for i in range(3):
        for j in range(4):
            h = 1e-2 if j == 3 else 1e-4
            mp, mm = cands[k].copy(), cands[k].copy()
            mp[i, j] += h
            mm[i, j] -= h
            fd[i, j] = (loss(mp) - loss(mm)) / (2 * h)
    for j, sl in (("rotation", np.s_[:3, :3]), ("translation", np.s_[:3, 3])):
        err = np.abs(g[sl] - fd[sl]).max() / np.abs(fd[sl]).max()
        print(f"[full-size] PoU {j} entries: rel err vs FD {err:.2e}")
        assert err <= 2e-3, (j, g[sl], fd[sl])
    e16 = np.abs(grads["m16"] - grads["m32"]).max() / np.abs(grads["m32"]).max()
    print(f"[full-size] PoU bf16 vs f32 gradient: {e16:.2e}")
    assert e16 <= 5e-2


# ---------------------------------------------------------------------------
# ragged multistart: 128^3 / 192^3 / 256^3, 512 poses split 96 / 160 / 256
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def ragged():
    from fvr import DifferentiableFourierVolumeRenderer
    from fvr.multistart import autotune
    from fvr.phantom import ellipsoid_phantom, random_poses
    from fvr.sncc import BiplanarEvaluator, RaggedBiplanarEvaluator
    sizes, counts = (128, 192, 256), (96, 160, 256)
    segs, truths, starts, oracle = [], [], [], []
    for k, (n, c) in enumerate(zip(sizes, counts)):
        vol = ellipsoid_phantom(n, n_ellipsoids=12, seed=1234 + k)
        rng = np.random.default_rng(2345 + k)
        tp = np.concatenate([rng.uniform(-5, 5, 3), rng.uniform(-5, 5, 3)])
        print(f"[full-size] ragged: oracle spectrum {2 * n}^3", flush=True)
        f3 = O.make_spectrum(vol)
        T = _oracle_targets(f3, vol.shape, tp)
        r = DifferentiableFourierVolumeRenderer(vol, pad_factor=2.0, storage="fp16", device=DEV)
        segs.append(BiplanarEvaluator(r.packed, *(torch.from_numpy(t.astype(np.complex64)).to(DEV) for t in T)))
        truths.append(tp)
        starts.append(tp + random_poses(c, 8.0, 8.0, seed=3456 + k))
        oracle.append((f3, [_c64(t) for t in T]))
    rag = RaggedBiplanarEvaluator(segs, counts)
    P0 = torch.from_numpy(np.ascontiguousarray(np.vstack(starts))).to(DEV)
    plan, _ = autotune(rag, P0, reps=2)
    return dict(rag=rag, segs=segs, counts=counts, sizes=sizes, truths=truths, P0=P0, oracle=oracle, plan=plan)


def test_ragged_512_poses_match_oracle(ragged):
    rag, P0 = ragged["rag"], ragged["P0"]
    L, G = (x.cpu().numpy().copy() for x in rag.evaluate_device(P0))
    poses = P0.cpu().numpy()
    off = 0
    for k, c in enumerate(ragged["counts"]):
        f3, (ta, tl) = ragged["oracle"][k]
        for j in (off, off + c // 2, off + c - 1):
            Lr, gr = O.biplanar_loss_and_grad_analytic(f3, poses[j], ta, tl, fold=True)
            assert abs(L[j] - Lr) <= 1e-4, (k, j, L[j], Lr)
            assert np.abs(G[j] - gr).max() <= 1e-3 * np.abs(gr).max(), (k, j, G[j], gr)
        off += c


def test_ragged_100_step_graph_loop_equals_eager_and_recovers_poses(ragged):
    from fvr.multistart import AdamLoop
    from fvr.registration_3d2d import compute_mtre_6dof
    rag, P0 = ragged["rag"], ragged["P0"]
    n = P0.shape[0]
    eager = AdamLoop(rag, n, 100, graph=False, tol=0.0)
    bp, bl, last = (t.cpu().numpy().copy() for t in eager.run(P0))
    loop = AdamLoop(rag, n, 100, graph=True, tol=0.0)
    for rep in range(2):
        gp, gl, glast = (t.cpu().numpy().copy() for t in loop.run(P0))
        np.testing.assert_array_equal(gp, bp, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(gl, bl, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(glast, last, err_msg=f"replay {rep}")
        rag.evaluate_device(P0)  # unrelated eager work between replays
        torch.cuda.synchronize()
    off = 0
    for k, (n_k, c) in enumerate(zip(ragged["sizes"], ragged["counts"])):
        kb = off + int(np.argmin(bl[off:off + c]))
        mtre = compute_mtre_6dof(ragged["truths"][k], bp[kb], (n_k,) * 3)
        print(f"[full-size] ragged {n_k}^3: best loss {bl[kb]:.3e}, mTRE {mtre:.4f} mm")
        assert mtre < 0.5, (n_k, mtre)
        off += c
