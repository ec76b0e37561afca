"""GPU parity of edge_dominant_ncc (cone_beam_fvr.py:71-112) on the HIP kernels against
the FP64 oracle (oracle/cone_beam_np.py: edge_dominant_ncc, pinned by scipy / torch
stencils, the reference's known answers, invariants and finite differences).

Tolerances (float32 images and stencils, float64 moments): values within 1e-5,
gradients within 1e-4 of max|grad|; through the PoU projection, the pose gradient
within 2e-3 of max|grad| of FP64 central differences of the oracle chain.
"""

import numpy as np
import pytest
import torch

from oracle import cone_beam_np as C

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _images(B, H, W, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((B, 1, H, W)).cumsum(-1) * 0.1
    t = rng.standard_normal((B, 1, H, W)).cumsum(-2) * 0.1 + 0.3 * p
    return p.astype(np.float32), t.astype(np.float32)


@pytest.mark.parametrize("H,W,crop,Bt", [(100, 104, None, 3), (100, 104, None, 1), (70, 90, None, 3),
                                         (96, 96, 0, 1), (130, 97, 7, 3), (40, 41, None, 1)])
def test_edge_ncc_values_and_gradients_match_oracle(H, W, crop, Bt):
    from fvr import edge_dominant_ncc
    p, t = _images(3, H, W, H + W)
    t = t[:Bt]
    P = torch.from_numpy(p).to(DEV).requires_grad_(True)
    T = torch.from_numpy(t).to(DEV).requires_grad_(True)
    val = edge_dominant_ncc(P, T, border_crop=crop, reduction="none")
    assert val.shape == (3,) and val.dtype == torch.float32
    w = torch.tensor([0.7, -1.3, 0.4], device=DEV)
    (val * w).sum().backward()
    ref, g = C.edge_dominant_ncc(p.astype(np.float64), t.astype(np.float64), border_crop=crop, reduction="none",
                                 with_grad=True)
    assert np.abs(val.detach().double().cpu().numpy() - ref).max() <= 1e-5, (val, ref)
    wn = w.double().cpu().numpy()[:, None, None, None]
    gp = P.grad.double().cpu().numpy()
    assert np.abs(gp - g * wn).max() <= 1e-4 * np.abs(g * wn).max()
    # the target's gradient: the value is symmetric, so it is the oracle's with the roles swapped
    tt = np.broadcast_to(t.astype(np.float64), p.shape)
    _, gt = C.edge_dominant_ncc(tt, p.astype(np.float64), border_crop=crop, reduction="none", with_grad=True)
    gt = (gt * wn).sum(0, keepdims=True) if Bt == 1 else gt * wn
    gT = T.grad.double().cpu().numpy()
    assert gT.shape == gt.shape and np.abs(gT - gt).max() <= 1e-4 * np.abs(gt).max()
    # the scalar reduction and a deterministic second backward
    m = edge_dominant_ncc(P, T, border_crop=crop)
    assert abs(float(m) - ref.mean()) <= 1e-5
    P2 = P.detach().clone().requires_grad_(True)
    (edge_dominant_ncc(P2, T.detach(), border_crop=crop, reduction="none") * w).sum().backward()
    assert torch.equal(P2.grad, P.grad)


def test_edge_ncc_reference_known_answers():
    """The reference's vignetting case (tests/test_cone_beam_fvr.py:95-107) on the HIP path."""
    from fvr import edge_dominant_ncc
    H = 96
    yy, xx = torch.meshgrid(torch.linspace(-1, 1, H), torch.linspace(-1, 1, H), indexing="ij")
    img = torch.exp(-((xx - 0.1) ** 2 + (yy + 0.05) ** 2) / 0.04)
    img = img + 0.5 * ((xx ** 2 + yy ** 2) < 0.25).float()
    ref = img[None, None].to(DEV)
    vignette = ref.clone() + 0.8 * xx[None, None].to(DEV)
    vignette[:, :, :6, :] = -10.0
    vignette[:, :, -6:, :] = -10.0
    shifted = torch.roll(ref, shifts=(4, 4), dims=(-2, -1))
    assert float(edge_dominant_ncc(ref, ref)) > 0.999
    assert float(edge_dominant_ncc(ref, vignette)) > 0.98
    assert float(edge_dominant_ncc(shifted, vignette)) < 0.50
    for a, b in ((ref, vignette), (shifted, vignette)):
        exp = C.edge_dominant_ncc(a.double().cpu().numpy(), b.double().cpu().numpy())
        assert abs(float(edge_dominant_ncc(a, b)) - exp) <= 1e-5


def test_edge_ncc_rejects_bad_shapes():
    from fvr import edge_dominant_ncc
    a = torch.zeros((2, 1, 32, 32), device=DEV)
    for b in (torch.zeros((3, 1, 32, 32), device=DEV), torch.zeros((2, 1, 32, 31), device=DEV),
              torch.zeros((2, 2, 32, 32), device=DEV)):
        with pytest.raises(ValueError):
            edge_dominant_ncc(a, b)
    with pytest.raises(ValueError):
        edge_dominant_ncc(a, a, border_crop=16)


def test_edge_ncc_through_the_pinhole_projection_matches_oracle_fd():
    """1 - edge_dominant_ncc(forward_pinhole(M), target) on a 96^2 detector (the blurred
    term on): the pose gradient (fused PoU adjoint after the edge adjoint) against FP64
    central differences of the oracle's chain, on the translation entries of two poses."""
    from fvr import PerspectiveSlabFVR, edge_dominant_ncc
    from scipy.spatial.transform import Rotation
    n = 24
    vol = np.zeros((n, n, n), dtype=np.float32)
    zz, yy, xx = np.ogrid[:n, :n, :n]
    vol[((zz - 11) ** 2 + (yy - 13) ** 2 + (xx - 12) ** 2) < 7 ** 2] = 0.08
    vol[9:15, 5:9, 14:19] += 0.05
    m = PerspectiveSlabFVR.from_volume(vol, vox_mm=2.0, grid_blocks=(2, 2, 2), block_res=16)
    o = C.from_volume(vol.astype(np.float64), 2.0, (2, 2, 2), 16)
    H = W = 96
    k_inv = np.array([[1.0 / 600.0, 0.0, -W / 1200.0], [0.0, 1.0 / 600.0, -H / 1200.0], [0.0, 0.0, 1.0]])
    lat = m.make_pinhole_lattice(torch.from_numpy(k_inv).float(), 600.0, H, W)
    lat_np = C.make_pinhole_lattice(k_inv, 600.0, H, W)
    mats = np.tile(np.eye(4), (2, 1, 1))
    mats[:, :3, :3] = Rotation.from_rotvec([[0.1, -0.05, 0.08], [-0.12, 0.06, -0.03]]).as_matrix()
    mats[:, :3, 3] = [[3.0, -2.0, -300.0], [-1.5, 2.5, -305.0]]
    true = mats[:1].copy()
    true[0, :3, 3] += [1.0, -1.0, 2.0]
    with torch.no_grad():
        target = m.forward_pinhole(torch.from_numpy(true).float(), lat)
    M = torch.from_numpy(mats).float().to(DEV).requires_grad_(True)
    loss = 1.0 - edge_dominant_ncc(m.forward_pinhole(M, lat), target)
    loss.backward()
    g = M.grad.double().cpu().numpy()
    tgt = target.double().cpu().numpy()

    def loss64(Mk):
        return 1.0 - C.edge_dominant_ncc(C.forward_pinhole(o, Mk, lat_np), tgt)

    assert abs(float(loss) - loss64(mats)) <= 1e-4
    h = 1e-3
    fd = np.zeros((2, 3))
    for k in range(2):
        for i in range(3):
            mp, mm = mats.copy(), mats.copy()
            mp[k, i, 3] += h
            mm[k, i, 3] -= h
            fd[k, i] = (loss64(mp) - loss64(mm)) / (2 * h)
    assert np.abs(g[:, :3, 3] - fd).max() <= 2e-3 * np.abs(fd).max(), (g[:, :3, 3], fd)


def _ncc_torch64(p, t):
    p0 = p - p.mean(dim=tuple(range(1, p.dim())), keepdim=True)
    t0 = t - t.mean(dim=tuple(range(1, t.dim())), keepdim=True)
    dims = tuple(range(1, p.dim()))
    return (p0 * t0).sum(dim=dims) / (p0.square().sum(dim=dims).sqrt() * t0.square().sum(dim=dims).sqrt() + 1e-8)


@pytest.mark.parametrize("B,shape,Bt", [(3, (1, 50, 61), 3), (3, (1, 50, 61), 1), (2, (1, 100, 104), 1),
                                        (4, (7, 9), 4), (64, (1, 256, 256), 1)])
def test_image_ncc_values_and_gradients(B, shape, Bt):
    """image_ncc (the reference test's NCC, tests/test_cone_beam_fvr.py:24-27) against the
    FP64 oracle's ncc_loss per image and float64 torch autograd of the same formula for both
    arguments: float32 images (rows of K % 4 != 0 take the scalar loads, K > 4096 several
    chunks); values within 1e-6, gradients within 1e-5 of max|grad|."""
    from fvr import image_ncc
    rng = np.random.default_rng(B + Bt + shape[-1])
    p = (rng.standard_normal((B,) + shape).cumsum(-1) * 0.1).astype(np.float32)
    t = (rng.standard_normal((Bt,) + shape).cumsum(-2) * 0.1 + 0.3 * p[:Bt]).astype(np.float32)
    P = torch.from_numpy(p).to(DEV).requires_grad_(True)
    T = torch.from_numpy(t).to(DEV).requires_grad_(True)
    val = image_ncc(P, T)
    assert val.shape == (B,) and val.dtype == torch.float32
    w = torch.from_numpy(rng.uniform(-1.0, 1.0, B).astype(np.float32)).to(DEV)
    (val * w).sum().backward()
    ref = np.array([1.0 - C.ncc_loss(p[n].astype(np.float64), t[n if Bt > 1 else 0].astype(np.float64))
                    for n in range(B)])
    assert np.abs(val.detach().double().cpu().numpy() - ref).max() <= 1e-6
    P64 = torch.from_numpy(p.astype(np.float64)).requires_grad_(True)
    T64 = torch.from_numpy(t.astype(np.float64)).requires_grad_(True)
    (_ncc_torch64(P64, T64.expand_as(P64)) * w.double().cpu()).sum().backward()
    for got, exp in ((P.grad, P64.grad), (T.grad, T64.grad)):
        e = exp.numpy()
        assert got.shape == exp.shape
        assert np.abs(got.double().cpu().numpy() - e).max() <= 1e-5 * np.abs(e).max()


def test_image_ncc_constant_image_and_shape_checks():
    """A constant image has NCC 0 (the 1e-8 of the denominator) and a finite gradient (no
    |p0| term); mismatched shapes raise before any launch."""
    from fvr import image_ncc
    p = torch.full((2, 1, 16, 16), 0.5, device=DEV, requires_grad=True)
    t = torch.randn((1, 1, 16, 16), generator=torch.Generator().manual_seed(2)).to(DEV)
    val = image_ncc(p, t)
    assert torch.all(val == 0)
    val.sum().backward()
    g = p.grad.double().cpu().numpy()
    td = t.double().cpu().numpy()
    e = (td - td.mean()) / 1e-8
    assert np.all(np.isfinite(g)) and np.abs(g - e).max() <= 1e-5 * np.abs(e).max()
    with pytest.raises(ValueError):
        image_ncc(p, torch.zeros((1, 1, 16, 15), device=DEV))
    with pytest.raises(ValueError):
        image_ncc(p, torch.zeros((3, 1, 16, 16), device=DEV))


def test_edge_ncc_target_cache_follows_the_target_tensor():
    """The target's channel images and moments are kept for the target tensor last seen: a
    second call equals the first bit for bit, and after the target is changed in place (its
    version counter moves) the value equals that of a fresh copy of the changed target."""
    from fvr import edge_dominant_ncc
    p, t = _images(4, 100, 104, 5)
    P = torch.from_numpy(p).to(DEV)
    T = torch.from_numpy(t[:1]).to(DEV)
    v1 = edge_dominant_ncc(P, T, reduction="none")
    v2 = edge_dominant_ncc(P, T, reduction="none")
    assert torch.equal(v1, v2)
    T.mul_(0.5).add_(0.01 * P[:1])
    v3 = edge_dominant_ncc(P, T, reduction="none")
    v4 = edge_dominant_ncc(P, T.clone(), reduction="none")
    assert torch.equal(v3, v4) and not torch.equal(v1, v3)
