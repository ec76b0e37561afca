"""FP64 NumPy oracle for the reference's PoU-FVR perspective path (convention P).

TEST INFRASTRUCTURE ONLY -- imported by ``tests/`` (never by the product
package).  Each function restates one piece of ``src/fvr/cone_beam_fvr.py`` or
``src/fvr/triton_slice.py`` of pareelamre/diff-fvr (cited per function) in
float64.  The per-corner trilinear slice is ``torch_engine_np.slice_per_corner``.

How each function is pinned (``tests/test_oracle_pou.py``):

* ``make_pou_1d``      -- the windows sum to 1 (the reference's
                          ``test_pou_windows_sum_to_unity``, atol 1e-5; here 1e-12).
* ``resample_linear``  -- equals ``torch.nn.functional.interpolate``
                          (trilinear, align_corners=False) in float64.
* ``build_blocks``     -- every block cube equals the PoU-weighted volume
                          sampled (scipy ``map_coordinates``, trilinear) at the
                          world points its samples stand for (centre + (q - P//2)
                          ds), times the sinc^2 window and ds, to 1e-12 inside
                          the block support, with and without resampling, on a
                          non-cubic volume and block grid (a half-voxel centre
                          error fails); the cubes, their window and ds divided
                          out, re-sum to the volume; the reference's known
                          answers: centre offset 1.25 mm between pixel- and
                          grid-centre origins at 2.5 mm
                          (``test_grid_center_vs_pixel_center_origin``), 64
                          blocks for the isotropic preset
                          (``test_slab_presets_and_parity_offsets``).
* ``sinc2_precomp_1d`` -- the closed form from sin(pi c) / (pi c), the 0.05 clamp.
* ``slab_images``      -- Fourier-slice theorem: at an axis-aligned detector the
                          slab image is the ray sum of the padded, weighted cube
                          (samples hit lattice points exactly).
* ``warp_sum``         -- the explicit 4-neighbour stencil equals
                          ``grid_sample`` (float64, align_corners=True, zeros),
                          the relation of ``test_pou_pinhole_warp_sum_analytical_parity``.
* ``correlate3`` (+adjoint) -- ``scipy.ndimage.correlate`` (zero padding) and
                          torch ``conv2d`` / ``avg_pool2d`` in float64; adjoints by
                          <K p, g> = <p, K^T g>.
* ``edge_dominant_ncc`` -- the reference's vignetting known answers
                          (``test_edge_dominant_ncc_vignetting_invariance``);
                          NCC(p, p) = 1, symmetry, invariance to a t + b (a > 0),
                          sign flip, batch mean, the blurred term's weighting and
                          its W >= 96 switch, rows inside the crop not counting;
                          gradient by FP64 directional differences.
* ``forward_pinhole`` / ``project_rays_batched`` -- SE(3)-invariant lattice path
                          equals the explicit-ray path
                          (``test_invariant_pinhole_matches_explicit_ray_tracer``);
                          batched equals single; front/back magnification
                          (``test_perspective_differential_magnification_front_vs_back``);
                          brute-force line integrals of a smooth phantom along the
                          pinhole rays (source 3 m away): centroids within 0.3 px,
                          values within 12 % of the peak (the method's own
                          approximation), a 1 mm centre error fails.
"""

from __future__ import annotations

import math
from typing import Dict, List, Optional, Tuple

import numpy as np

from .torch_engine_np import slice_per_corner

SLAB_PRESETS = {
    "fast_coarse": {"grid_blocks": (2, 2, 2), "block_res": 64},
    "balanced": {"grid_blocks": (2, 4, 2), "block_res": 80},
    "isotropic": {"grid_blocks": (4, 4, 4), "block_res": 64},
}


def make_pou_1d(length: int, n_blocks: int, overlap: float) -> Tuple[List[np.ndarray], float]:
    """Raised-cosine partition-of-unity windows (cone_beam_fvr.py:42-61)."""
    if n_blocks <= 1:
        return [np.ones(length)], float(length)
    step = length / float(n_blocks)
    ov = max(2.0, step * overlap)
    x = np.arange(length) + 0.5
    out = []
    for i in range(n_blocks):
        w = np.ones(length)
        if i > 0:
            t = np.clip((x - (i * step - ov)) / (2 * ov), 0.0, 1.0)
            w = w * 0.5 * (1.0 - np.cos(np.pi * t))
        if i < n_blocks - 1:
            t = np.clip((x - ((i + 1) * step - ov)) / (2 * ov), 0.0, 1.0)
            w = w * 0.5 * (1.0 + np.cos(np.pi * t))
        out.append(w)
    return out, step


def sinc2_precomp_1d(n: int, p: int) -> np.ndarray:
    """1 / max(sinc^2((i - (n-1)/2) / p), 0.05) (cone_beam_fvr.py:64-68)."""
    c = (np.arange(n) - (n - 1) / 2.0) / float(p)
    return 1.0 / np.maximum(np.sinc(c) ** 2, 0.05)


def resample_linear(cube: np.ndarray, r: int) -> np.ndarray:
    """Separable linear resampling of an (M, M, M) cube to (r, r, r) with PyTorch's
    ``interpolate(mode='trilinear', align_corners=False)`` source-index rule."""
    out = cube
    for axis in range(3):
        m = out.shape[axis]
        src = np.maximum((np.arange(r) + 0.5) * (m / r) - 0.5, 0.0)
        i0 = np.floor(src).astype(np.int64)
        i0 = np.minimum(i0, m - 1)
        i1 = np.minimum(i0 + 1, m - 1)
        lam = src - i0
        a = np.take(out, i0, axis=axis)
        b = np.take(out, i1, axis=axis)
        shape = [1, 1, 1]
        shape[axis] = r
        lam = lam.reshape(shape)
        out = a * (1.0 - lam) + b * lam
    return out


def build_blocks(vol: np.ndarray, vox_mm: float = 1.0, origin=(0.0, 0.0, 0.0), axis_signs=(1.0, 1.0, 1.0),
                 grid_blocks=(2, 2, 2), block_res: int = 80, pad_mult: int = 2, overlap: float = 0.25,
                 energy_thresh: float = 1e-4) -> Dict:
    """PoU decomposition + per-block centred spectra (cone_beam_fvr.py:172-303).

    Returns ``F_blk`` (B, P, P, P) complex128, ``centers`` (B, 3) world mm, ``ds``, ``P``,
    ``B`` and the spatial ``cubes`` (B, P, P, P) before the FFT.
    """
    vol = np.asarray(vol, dtype=np.float64)
    n0, n1, n2 = vol.shape
    P = block_res * pad_mult
    k0, k1, k2 = grid_blocks
    wl = [make_pou_1d(n, k, overlap) for n, k in zip((n0, n1, n2), grid_blocks)]
    ov = [max(2, int(round(s * overlap))) if k > 1 else 0 for (_, s), k in zip(wl, grid_blocks)]
    m = [int(np.ceil(s + 2 * o)) for (_, s), o in zip(wl, ov)]
    M = max(m) + (1 if max(grid_blocks) > 1 else 0)
    ds = (M * vox_mm) / float(block_res)
    ws = sinc2_precomp_1d(block_res, P)
    w3 = ws[:, None, None] * ws[None, :, None] * ws[None, None, :]
    pad_b = (P - block_res) // 2
    half_vox_mm = ((P // 2) - (pad_b + (block_res - 1) * 0.5)) * ds

    def support(w):
        nz = np.nonzero(w > 1e-6)[0]
        return int(nz[0]), int(nz[-1]) + 1

    cubes, centers = [], []
    for w0 in wl[0][0]:
        a0, b0 = support(w0)
        for w1 in wl[1][0]:
            a1, b1 = support(w1)
            for w2 in wl[2][0]:
                a2, b2 = support(w2)
                sub = vol[a0:b0, a1:b1, a2:b2] * (w0[a0:b0, None, None] * w1[None, a1:b1, None]
                                                  * w2[None, None, a2:b2])
                if float(sub.sum()) <= energy_thresh:
                    continue
                l = sub.shape
                mm = [(M - li) // 2 for li in l]
                cube = np.zeros((M, M, M))
                cube[mm[0]:mm[0] + l[0], mm[1]:mm[1] + l[1], mm[2]:mm[2] + l[2]] = sub
                cidx = [a + (M - 1) * 0.5 - q for a, q in zip((a0, a1, a2), mm)]
                centers.append([origin[k] + axis_signs[k] * cidx[k] * vox_mm + axis_signs[k] * half_vox_mm
                                for k in range(3)])
                cube_r = (resample_linear(cube, block_res) if M != block_res else cube) * ds
                padded = np.zeros((P, P, P))
                sl = slice(pad_b, pad_b + block_res)
                padded[sl, sl, sl] = cube_r * w3
                cubes.append(padded)
    if not cubes:
        raise ValueError("no block above the energy threshold")
    cubes = np.stack(cubes)
    F = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(cubes, axes=(-3, -2, -1)), axes=(-3, -2, -1)),
                        axes=(-3, -2, -1))
    return {"F_blk": F, "centers": np.array(centers), "ds": ds, "P": P, "B": len(cubes), "cubes": cubes,
            "vox_mm": vox_mm}


def from_volume(vol, vox_mm=1.0, grid_blocks=None, block_res=None, pad_mult=2, overlap=0.25,
                center_mode="pixel_center", preset=None, **kw) -> Dict:
    """PerspectiveSlabFVR.from_volume (cone_beam_fvr.py:332-374)."""
    if preset is not None:
        if preset not in SLAB_PRESETS:
            raise ValueError(preset)
        grid_blocks = grid_blocks or SLAB_PRESETS[preset]["grid_blocks"]
        block_res = block_res or SLAB_PRESETS[preset]["block_res"]
    grid_blocks = grid_blocks or (2, 2, 2)
    block_res = block_res or 64
    nz, ny, nx = np.asarray(vol).shape
    v = float(vox_mm)
    if center_mode == "pixel_center":
        origin = (-0.5 * (nz - 1) * v, -0.5 * (ny - 1) * v, -0.5 * (nx - 1) * v)
    elif center_mode == "grid_center":
        origin = (-0.5 * nz * v, -0.5 * ny * v, -0.5 * nx * v)
    else:
        raise ValueError(center_mode)
    return build_blocks(vol, v, origin, (1.0, 1.0, 1.0), grid_blocks, block_res, pad_mult, overlap, **kw)


# ---------------------------------------------------------------------------
# projection (cone_beam_fvr.py:480-545, 576-596, 670-728)
# ---------------------------------------------------------------------------
def slab_images(model: Dict, u_det: np.ndarray, v_det: np.ndarray) -> np.ndarray:
    """(B, N, P, P) real slab images: central slice along (u_det, v_det) + centred iFFT2
    (cone_beam_fvr.py:480-505; u_det, v_det in world (axis0, axis1, axis2) order)."""
    F, P = model["F_blk"], model["P"]
    u = np.asarray(u_det, np.float64).reshape(-1, 3)[:, ::-1]  # flip -> (x, y, z)
    v = np.asarray(v_det, np.float64).reshape(-1, 3)[:, ::-1]
    k = np.arange(P) - P // 2  # fftshift(fftfreq(P, d)) * P * d: integer bins
    c = float(P // 2)
    out = np.empty((F.shape[0], len(u), P, P))
    for n in range(len(u)):
        # rows follow FY (v), columns follow FX (u)
        p = c + k[None, :, None] * u[n][None, None, :] + k[:, None, None] * v[n][None, None, :]
        for b in range(F.shape[0]):
            s = slice_per_corner(F[b].real, p) + 1j * slice_per_corner(F[b].imag, p)
            out[b, n] = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(s))).real
    return out


def block_offsets(centers, srcs, u_det, v_det, n_0):
    """(d_b, cu, cv) each (B, N) (cone_beam_fvr.py:507-521)."""
    diff = np.asarray(centers)[:, None, :] - np.asarray(srcs)[None, :, :]
    return (np.einsum("bnk,nk->bn", diff, n_0), np.einsum("bnk,nk->bn", diff, u_det),
            np.einsum("bnk,nk->bn", diff, v_det))


def extract_detector_basis(srcs, trg_pts):
    """Orthonormal (u_det, v_det, n_0) from a detector lattice (cone_beam_fvr.py:576-596)."""
    srcs = np.asarray(srcs, np.float64).reshape(-1, 3)
    trg = np.asarray(trg_pts, np.float64).reshape(len(srcs), *np.asarray(trg_pts).shape[-3:])
    c_det = trg.mean(axis=(1, 2))
    u = trg[:, 0, -1] - trg[:, 0, 0]
    u = u / (np.linalg.norm(u, axis=-1, keepdims=True) + 1e-12)
    v = trg[:, -1, 0] - trg[:, 0, 0]
    v = v / (np.linalg.norm(v, axis=-1, keepdims=True) + 1e-12)
    n = np.cross(v, u)
    n = n / (np.linalg.norm(n, axis=-1, keepdims=True) + 1e-12)
    flip = np.sum((c_det - srcs) * n, axis=-1, keepdims=True) < 0
    return u, v, np.where(flip, -n, n)


def make_ray_cache(src0, trg0, u0, v0, n0):
    """SE(3)-invariant (rx, ry, inv_cos) of a camera-frame lattice (cone_beam_fvr.py:524-545)."""
    rays = np.asarray(trg0, np.float64) - np.asarray(src0, np.float64).reshape(1, 1, 3)
    r = rays / np.sqrt(np.maximum(np.sum(rays * rays, axis=-1, keepdims=True), 1e-12))
    cos0 = np.maximum(np.sum(r * n0, axis=-1), 0.1)
    ic = 1.0 / cos0
    return np.sum(r * u0, axis=-1) * ic, np.sum(r * v0, axis=-1) * ic, ic


def make_pinhole_lattice(k_inv, sdd_mm, height, width):
    """Camera-frame pinhole lattice + ray cache (cone_beam_fvr.py:706-728)."""
    k = np.asarray(k_inv, np.float64).reshape(3, 3)
    cols = np.arange(width) + 0.5
    rows = np.arange(height) + 0.5
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    pix = np.stack([cc, rr, np.ones_like(cc)], axis=-1)
    dirs = np.einsum("ij,hwj->hwi", k, pix)
    trg0 = dirs * (float(sdd_mm) / np.maximum(dirs[..., 2:3], 1e-6))
    src0 = np.zeros(3)
    u, v, n = extract_detector_basis(src0[None], trg0[None])
    u, v, n = u[0], v[0], n[0]
    rx, ry, ic = make_ray_cache(src0, trg0, u, v, n)
    return {"src0": src0, "trg0": trg0, "u0": u, "v0": v, "n0": n, "rx": rx, "ry": ry, "ic": ic}


def warp_sum(imgs, d_b, cu, cv, rx, ry, ic, inv_ds, c_idx):
    """sum_b bilinear(img_bn, p) * inv_cos with per-corner zero padding
    (triton_slice.py:349-407 / 748-784); ray arrays (H, W) shared or (N, H, W) per pose."""
    imgs = np.asarray(imgs, np.float64)
    B, N, P, _ = imgs.shape
    rx, ry, ic = (np.broadcast_to(np.asarray(a, np.float64), (N,) + np.asarray(a).shape[-2:]) for a in (rx, ry, ic))
    H, W = rx.shape[-2:]
    out = np.zeros((N, H, W))
    for n in range(N):
        for b in range(B):
            px = (d_b[b, n] * rx[n] - cu[b, n]) * inv_ds + c_idx
            py = (d_b[b, n] * ry[n] - cv[b, n]) * inv_ds + c_idx
            x0, y0 = np.floor(px), np.floor(py)
            tx, ty = px - x0, py - y0
            ix, iy = x0.astype(np.int64), y0.astype(np.int64)
            acc = 0.0
            for dx, dy, w in ((0, 0, (1 - tx) * (1 - ty)), (1, 0, tx * (1 - ty)), (0, 1, (1 - tx) * ty),
                              (1, 1, tx * ty)):
                gx, gy = ix + dx, iy + dy
                ok = (gx >= 0) & (gx < P) & (gy >= 0) & (gy < P)
                acc = acc + np.where(ok, imgs[b, n][np.clip(gy, 0, P - 1), np.clip(gx, 0, P - 1)], 0.0) * w
            out[n] += acc
        out[n] *= ic[n]
    return out


def forward_pinhole(model: Dict, rt_inv: np.ndarray, lattice: Dict) -> np.ndarray:
    """(N, 1, H, W) projections for SE(3) matrices ``rt_inv`` (cone_beam_fvr.py:670-688, 730-740)."""
    M = np.asarray(rt_inv, np.float64).reshape(-1, 4, 4)
    R, t = M[:, :3, :3], M[:, :3, 3]
    srcs = np.einsum("nij,j->ni", R, lattice["src0"]) + t
    u = np.einsum("nij,j->ni", R, lattice["u0"])
    v = np.einsum("nij,j->ni", R, lattice["v0"])
    n0 = np.einsum("nij,j->ni", R, lattice["n0"])
    imgs = slab_images(model, u, v)
    d_b, cu, cv = block_offsets(model["centers"], srcs, u, v, n0)
    out = warp_sum(imgs, d_b, cu, cv, lattice["rx"], lattice["ry"], lattice["ic"], 1.0 / model["ds"],
                   float(model["P"] // 2))
    return out[:, None]


def project_rays_batched(model: Dict, srcs, trg_pts, u_det=None, v_det=None, n_0=None) -> np.ndarray:
    """Explicit per-pose ray path (cone_beam_fvr.py:423-478, 598-616): (N, H, W)."""
    srcs = np.asarray(srcs, np.float64).reshape(-1, 3)
    trg = np.asarray(trg_pts, np.float64)
    u0, v0, n0 = extract_detector_basis(srcs, trg)
    u_det = u0 if u_det is None else np.asarray(u_det, np.float64).reshape(-1, 3)
    v_det = v0 if v_det is None else np.asarray(v_det, np.float64).reshape(-1, 3)
    n_0 = n0 if n_0 is None else np.asarray(n_0, np.float64).reshape(-1, 3)
    imgs = slab_images(model, u_det, v_det)
    d_b, cu, cv = block_offsets(model["centers"], srcs, u_det, v_det, n_0)
    rays = trg - srcs[:, None, None, :]
    r = rays / np.sqrt(np.maximum(np.sum(rays * rays, axis=-1, keepdims=True), 1e-12))
    ic = 1.0 / np.maximum(np.einsum("nhwk,nk->nhw", r, n_0), 0.1)
    rx = np.einsum("nhwk,nk->nhw", r, u_det) * ic
    ry = np.einsum("nhwk,nk->nhw", r, v_det) * ic
    return warp_sum(imgs, d_b, cu, cv, rx, ry, ic, 1.0 / model["ds"], float(model["P"] // 2))


def carm_geometry(yaw, pitch, roll, tx, ty, tz, sad_mm, sdd_mm, det_px, det_fov_mm, vox_mm, view="ap"):
    """Source, lattice and detector axes of forward_carm_view (cone_beam_fvr.py:761-832)."""
    y, p, r = (math.radians(a) for a in (yaw, pitch, roll))
    rz = np.array([[1, 0, 0], [0, math.cos(y), -math.sin(y)], [0, math.sin(y), math.cos(y)]])
    ry = np.array([[math.cos(r), 0, math.sin(r)], [0, 1, 0], [-math.sin(r), 0, math.cos(r)]])
    rx = np.array([[math.cos(p), -math.sin(p), 0], [math.sin(p), math.cos(p), 0], [0, 0, 1]])
    rot = rz @ ry @ rx
    lin = np.linspace(-0.5, 0.5, det_px) * float(det_fov_mm)
    vv, uu = np.meshgrid(lin, lin, indexing="ij")
    det_d = float(sdd_mm - sad_mm)
    if view == "ap":
        src0 = np.array([0.0, -float(sad_mm), 0.0])
        trg0 = np.stack([vv, np.full_like(uu, det_d), uu], axis=-1)
        u0, v0, n0 = np.array([0.0, 0, 1]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    else:
        src0 = np.array([0.0, 0.0, -float(sad_mm)])
        trg0 = np.stack([vv, uu, np.full_like(uu, det_d)], axis=-1)
        u0, v0, n0 = np.array([0.0, 1, 0]), np.array([1.0, 0, 0]), np.array([0.0, 0, 1])
    shift = np.array([tz, ty, tx]) * vox_mm
    return rot @ src0 - shift, np.einsum("ij,hwj->hwi", rot, trg0) - shift, rot @ u0, rot @ v0, rot @ n0


def forward_carm_view(model: Dict, yaw, pitch, roll, tx, ty, tz, sad_mm, sdd_mm, det_px, det_fov_mm,
                      view="ap") -> np.ndarray:
    src, trg, u, v, n = carm_geometry(yaw, pitch, roll, tx, ty, tz, sad_mm, sdd_mm, det_px, det_fov_mm,
                                      model["vox_mm"], view)
    return project_rays_batched(model, src[None], trg[None], u[None], v[None], n[None])[0]


def ncc_loss(a: np.ndarray, b: np.ndarray) -> float:
    """1 - NCC of zero-mean images (the reference test's loss, tests/test_cone_beam_fvr.py:24-27)."""
    a0, b0 = a - a.mean(), b - b.mean()
    return 1.0 - float(np.sum(a0 * b0) / (np.linalg.norm(a0) * np.linalg.norm(b0) + 1e-8))


# ---------------------------------------------------------------------------
# edge-dominant gradient NCC (cone_beam_fvr.py:23-39, 71-112)
# ---------------------------------------------------------------------------
SOBEL_X = np.array([[-1.0, 0.0, 1.0], [-2.0, 0.0, 2.0], [-1.0, 0.0, 1.0]])
SOBEL_Y = SOBEL_X.T.copy()


def correlate3(img: np.ndarray, k: np.ndarray) -> np.ndarray:
    """3 x 3 cross-correlation with zero padding over the last two axes (F.conv2d, padding=1)."""
    H, W = img.shape[-2:]
    p = np.pad(img, [(0, 0)] * (img.ndim - 2) + [(1, 1), (1, 1)])
    out = np.zeros(img.shape)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if k[dy + 1, dx + 1] != 0.0:
                out += k[dy + 1, dx + 1] * p[..., 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
    return out


def correlate3_adjoint(g: np.ndarray, k: np.ndarray) -> np.ndarray:
    """Adjoint of ``correlate3``: p_bar(q) = sum_k K(k) g(q - k) over q - k in the image."""
    H, W = g.shape[-2:]
    p = np.pad(g, [(0, 0)] * (g.ndim - 2) + [(1, 1), (1, 1)])
    out = np.zeros(g.shape)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if k[dy + 1, dx + 1] != 0.0:
                out += k[dy + 1, dx + 1] * p[..., 1 - dy:1 - dy + H, 1 - dx:1 - dx + W]
    return out


BOX3 = np.full((3, 3), 1.0 / 9.0)


def edge_dominant_ncc(pred, target, border_crop=None, multiscale_weight=0.25, reduction="mean", with_grad=False):
    """Gradient NCC of (B, C, H, W) images (cone_beam_fvr.py:71-112): Sobel x / y (zero padded),
    cropped by ``crop = max(4, W // 8)`` on every side (W = the last axis, as the reference
    reads it), NCC per channel with the mean over the crop removed and the denominator
    sqrt(max(Eu Ev, 1e-12)), the mean of x and y; for W >= 96 and a positive weight, (1 - w)
    times that plus w times the same on the 3 x 3 box-blurred images (zero padding, divisor
    9).  Channel mean, then the batch mean for reduction "mean".  ``with_grad``: also d(sum of
    the per-pose values) / d pred, or d(mean) / d pred for "mean" (exact adjoint)."""
    p = np.asarray(pred, np.float64)
    t = np.broadcast_to(np.asarray(target, np.float64), p.shape)
    W = p.shape[-1]
    crop = max(4, W // 8) if border_crop is None else int(border_crop)
    sl = (Ellipsis, slice(crop, -crop if crop > 0 else None), slice(crop, -crop if crop > 0 else None))

    def grad_ncc(a, b):
        val = 0.0
        grads = []
        for k in (SOBEL_X, SOBEL_Y):
            u, v = correlate3(a, k)[sl], correlate3(b, k)[sl]
            u0 = u - u.mean(axis=(-2, -1), keepdims=True)
            v0 = v - v.mean(axis=(-2, -1), keepdims=True)
            A = (u0 * v0).sum(axis=(-2, -1))
            Eu, Ev = (u0 * u0).sum(axis=(-2, -1)), (v0 * v0).sum(axis=(-2, -1))
            EE = Eu * Ev
            D = np.sqrt(np.maximum(EE, 1e-12))
            ncc = A / D
            val = val + 0.5 * ncc.mean(axis=1)
            if with_grad:  # d ncc / d u = v0 / D - (A / (D Eu)) u0 where Eu Ev > 1e-12, else v0 / D
                beta = np.where(EE > 1e-12, A / (D * np.where(Eu > 0, Eu, 1.0)), 0.0)
                gu = v0 / D[..., None, None] - beta[..., None, None] * u0
                full = np.zeros(a.shape)
                full[sl] = 0.5 * gu / a.shape[1]  # channel mean
                grads.append(correlate3_adjoint(full, k))
        return val, (sum(grads) if with_grad else None)

    fine, g_fine = grad_ncc(p, t)
    if multiscale_weight <= 0.0 or W < 96:
        val, g = fine, g_fine
    else:
        w = float(multiscale_weight)
        coarse, g_c = grad_ncc(correlate3(p, BOX3), correlate3(t, BOX3))
        val = (1.0 - w) * fine + w * coarse
        if with_grad:
            g = (1.0 - w) * g_fine + w * correlate3_adjoint(g_c, BOX3)
    out = float(np.mean(val)) if reduction == "mean" else val
    if not with_grad:
        return out
    return out, (g / p.shape[0] if reduction == "mean" else g)
