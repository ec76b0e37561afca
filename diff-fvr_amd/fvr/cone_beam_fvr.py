"""Drop-in ``PerspectiveSlabFVR`` (PoU-FVR perspective path) on HIP.

Restates ``src/fvr/cone_beam_fvr.py`` of pareelamre/diff-fvr.  The volume is
split into overlapping partition-of-unity blocks, each resampled to
``block_res``^3, sinc^2-precompensated, zero padded to P^3 and stored as its
centred spectrum (:172-303, built here in float64 on the device, then cast).
A projection evaluates, per (block, pose), the P x P central slice along the
detector axes, its centred inverse 2D FFT (the slab image), and the pinhole
warp that sums all blocks onto the detector with the 1/cos obliquity
(:423-545, 670-740, triton_slice.py:349-808).

HIP kernels (``csrc/fvr_pou.hip``): for power-of-two P <= 128 one launch
gathers each slice into LDS and inverse-transforms it there; the warp is one
launch over (pose, pixel); the backward is one fused launch per (block, pose)
(the warp adjoint -- a gather over each texel's detector footprint for the
separable rays of a zero-skew pinhole lattice, a scatter into an LDS fixed-point
image otherwise -- forward FFT in LDS, recomputed slice samples contracted with
their gradient) plus fixed-order block sums.
Other P use the generic slice operator, ``torch.fft`` and the warp kernels.
"""

from __future__ import annotations

import math
import weakref
from typing import Dict, Optional, Tuple, Union

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import _lib
from .triton_slice import _VOL_DTYPES, _PouWarpSum, _f32, _rays, fourier_slice_3d, pou_pinhole_warp_sum

SLAB_PRESETS: Dict[str, Dict[str, object]] = {
    "fast_coarse": {"grid_blocks": (2, 2, 2), "block_res": 64, "dtype": torch.bfloat16},
    "balanced": {"grid_blocks": (2, 4, 2), "block_res": 80, "dtype": torch.bfloat16},
    "isotropic": {"grid_blocks": (4, 4, 4), "block_res": 64, "dtype": torch.bfloat16},
}


def _make_pou_1d(length: int, n_blocks: int, overlap: float, device=None, dtype=torch.float64):
    """Raised-cosine partition-of-unity windows summing to 1 (cone_beam_fvr.py:42-61)."""
    if n_blocks <= 1:
        return [torch.ones(length, dtype=dtype, device=device)], float(length)
    step = length / float(n_blocks)
    ov = max(2.0, step * overlap)
    x = torch.arange(length, dtype=dtype, device=device) + 0.5
    out = []
    for i in range(n_blocks):
        w = torch.ones(length, dtype=dtype, device=device)
        if i > 0:
            t = torch.clamp((x - (i * step - ov)) / (2 * ov), 0.0, 1.0)
            w = w * 0.5 * (1.0 - torch.cos(math.pi * t))
        if i < n_blocks - 1:
            t = torch.clamp((x - ((i + 1) * step - ov)) / (2 * ov), 0.0, 1.0)
            w = w * 0.5 * (1.0 + torch.cos(math.pi * t))
        out.append(w)
    return out, step


def _sinc2_precomp_1d(n: int, p: int, device=None, dtype=torch.float64) -> torch.Tensor:
    """1 / max(sinc^2((i - (n-1)/2)/p), 0.05) (cone_beam_fvr.py:64-68)."""
    c = (torch.arange(n, dtype=dtype, device=device) - (n - 1) / 2.0) / float(p)
    return 1.0 / torch.clamp(torch.sinc(c) ** 2, min=0.05)


class _EdgeNCC(torch.autograd.Function):
    """Per-image edge-dominant gradient NCC (B,) of (B, H, W) float32 images against
    (Bt, H, W) targets, Bt in {1, B}, on the HIP kernels (``fvr_edge_*``); gradients for
    both arguments (the value is symmetric in them)."""

    # the reference image's channel images and moments, kept for the reference tensor last seen
    # (object, version, storage, crop, term): a registration loop against one target forms them once
    _ref_cache = None

    @staticmethod
    def _moments(img, ref, crop, multi, weight, ref_key=None):
        """(value (B,), coef, channel images of ref) of img against ref; ref_key (the caller's
        target tensor) lets the reference's channels and moments come from the cache."""
        L = _lib.lib()
        dev = img.device
        st = _lib.stream_ptr(dev)
        B, H, W = img.shape
        Bt = ref.shape[0]
        nch = 4 if multi else 2
        tiles = int(L.fvr_edge_tiles(H, W, crop))
        key = None
        if ref_key is not None:
            key = (id(ref_key), ref_key._version, ref_key.data_ptr(), tuple(ref_key.shape), ref_key.dtype, crop,
                   bool(multi), H, W)
            c = _EdgeNCC._ref_cache
            if c is not None and c[0] == key and c[1]() is ref_key:
                ch, tpart = c[2], c[3]
                key = None  # a hit: nothing to store
            else:
                c = None
        if ref_key is None or key is not None:
            ch = torch.empty((Bt, nch, H - 2 * crop, W - 2 * crop), device=dev, dtype=torch.float32)
            _lib.check(L.fvr_edge_channels(_lib.ptr(ref), Bt, H, W, crop, int(multi), _lib.ptr(ch), st),
                       "fvr_edge_channels")
            tpart = torch.empty((Bt, tiles, 12), device=dev, dtype=torch.float64)
            _lib.check(L.fvr_edge_stats(_lib.ptr(ref), Bt, H, W, crop, int(multi), _lib.ptr(ch), Bt,
                                        _lib.ptr(tpart), st), "fvr_edge_stats")
            if key is not None:
                _EdgeNCC._ref_cache = (key, weakref.ref(ref_key), ch, tpart)
        part = torch.empty((B, tiles, 12), device=dev, dtype=torch.float64)
        _lib.check(L.fvr_edge_stats(_lib.ptr(img), B, H, W, crop, int(multi), _lib.ptr(ch), Bt, _lib.ptr(part), st),
                   "fvr_edge_stats")
        val = torch.empty(B, device=dev, dtype=torch.float64)
        coef = torch.empty((B, 4, 4), device=dev, dtype=torch.float32)
        _lib.check(L.fvr_edge_combine(_lib.ptr(part), B, _lib.ptr(tpart), Bt, H, W, crop, int(multi), float(weight),
                                      _lib.ptr(val), _lib.ptr(coef), st), "fvr_edge_combine")
        return val, coef, ch

    @staticmethod
    def _grad(img, ch, Bt, coef, gval, crop, multi):
        L = _lib.lib()
        B, H, W = img.shape
        g = torch.empty((B, H, W), device=img.device, dtype=torch.float32)
        _lib.check(L.fvr_edge_bwd(_lib.ptr(img), B, H, W, crop, int(multi), _lib.ptr(ch), Bt, _lib.ptr(coef),
                                  _lib.ptr(gval), _lib.ptr(g), _lib.stream_ptr(img.device)), "fvr_edge_bwd")
        return g

    @staticmethod
    def forward(ctx, pred, target, crop, multi, weight, target_key=None):
        val, coef, ch = _EdgeNCC._moments(pred, target, crop, multi, weight, target_key)
        ctx.save_for_backward(pred, target, coef, ch)
        ctx.args = (crop, multi, weight)
        return val

    @staticmethod
    def backward(ctx, gval):
        pred, target, coef, ch = ctx.saved_tensors
        crop, multi, weight = ctx.args
        gv = gval.to(torch.float64).contiguous()
        gp = gt = None
        if ctx.needs_input_grad[0]:
            gp = _EdgeNCC._grad(pred, ch, target.shape[0], coef, gv, crop, multi)
        if ctx.needs_input_grad[1]:
            # the value is symmetric: d / d target is the same adjoint with the roles swapped,
            # per pose against that pose's prediction, summed over poses for a shared target
            tt = target.expand(pred.shape[0], -1, -1).contiguous()
            _, coef_t, ch_p = _EdgeNCC._moments(tt, pred, crop, multi, weight)
            gt = _EdgeNCC._grad(tt, ch_p, pred.shape[0], coef_t, gv, crop, multi)
            if target.shape[0] == 1 and pred.shape[0] > 1:
                gt = gt.sum(0, keepdim=True)
        return gp, gt, None, None, None, None


def edge_dominant_ncc(pred: torch.Tensor, target: torch.Tensor, border_crop: Optional[int] = None,
                      multiscale_weight: float = 0.25, reduction: str = "mean") -> torch.Tensor:
    """Gradient NCC over interior Sobel gradients, plus the 3 x 3-smoothed scale
    (cone_beam_fvr.py:71-112), on the HIP kernels ``fvr_edge_*``: pred (B, 1, H, W), target
    (B or 1, 1, H, W) on a ROCm device; crop = max(4, W // 8) unless given (W = the last
    axis, as the reference reads it); the blurred term for W >= 96 and a positive weight.
    reduction "mean" gives a scalar, anything else the (B,) per-image values.  Computed in
    float32 with float64 moments; the result and the gradients are in pred's dtype."""
    if pred.dim() != 4 or target.dim() != 4 or pred.shape[1] != 1 or target.shape[1] != 1:
        raise ValueError(f"edge_dominant_ncc: (B, 1, H, W) images expected, got {tuple(pred.shape)} and "
                         f"{tuple(target.shape)}")
    if target.shape[-2:] != pred.shape[-2:] or target.shape[0] not in (1, pred.shape[0]):
        raise ValueError(f"edge_dominant_ncc: target {tuple(target.shape)} does not match pred {tuple(pred.shape)}")
    _lib.require_device(pred, "edge_dominant_ncc")
    B, _, H, W = pred.shape
    crop = max(4, W // 8) if border_crop is None else int(border_crop)
    if crop < 0 or H - 2 * crop <= 0 or W - 2 * crop <= 0:
        raise ValueError(f"edge_dominant_ncc: crop {crop} leaves nothing of a {H} x {W} image")
    multi = multiscale_weight > 0.0 and W >= 96
    p = pred.to(torch.float32).reshape(B, H, W).contiguous()
    t = target.to(device=pred.device, dtype=torch.float32).reshape(target.shape[0], H, W).contiguous()
    # the target's channels and moments are cached for the target tensor object (its version
    # counter invalidates the entry when it is changed in place)
    key = target if not target.requires_grad else None
    val = _EdgeNCC.apply(p, t, crop, bool(multi), float(multiscale_weight), key).to(pred.dtype)
    return val.mean() if reduction == "mean" else val


class _ImageNCC(torch.autograd.Function):
    """Per-image zero-mean NCC (B,) of (B, K) float32 images against (Bt, K) targets, Bt in
    {1, B}, on the HIP kernels ``fvr_ncc_*``; gradients for both arguments."""

    @staticmethod
    def forward(ctx, pred, target):
        L = _lib.lib()
        B, K = pred.shape
        Bt = target.shape[0]
        dev = pred.device
        st = _lib.stream_ptr(dev)
        work = torch.empty(int(L.fvr_ncc_work_doubles(B, K)), device=dev, dtype=torch.float64)
        val = torch.empty(B, device=dev, dtype=torch.float64)
        coef = torch.empty((B, 8), device=dev, dtype=torch.float32)
        _lib.check(L.fvr_ncc_fwd(_lib.ptr(pred), B, _lib.ptr(target), Bt, K, _lib.ptr(work), _lib.ptr(val),
                                 _lib.ptr(coef), st), "fvr_ncc_fwd")
        ctx.save_for_backward(pred, target, coef)
        return val

    @staticmethod
    def backward(ctx, gval):
        pred, target, coef = ctx.saved_tensors
        L = _lib.lib()
        B, K = pred.shape
        Bt = target.shape[0]
        st = _lib.stream_ptr(pred.device)
        gv = gval.to(torch.float64).contiguous()
        gp = gt = None
        for wrt in (0, 1):
            if not ctx.needs_input_grad[wrt]:
                continue
            g = torch.empty((B, K), device=pred.device, dtype=torch.float32)
            _lib.check(L.fvr_ncc_bwd(_lib.ptr(pred), B, _lib.ptr(target), Bt, K, _lib.ptr(coef), _lib.ptr(gv), wrt,
                                     _lib.ptr(g), st), "fvr_ncc_bwd")
            if wrt == 0:
                gp = g
            else:
                gt = g.sum(0, keepdim=True) if Bt == 1 and B > 1 else g
        return gp, gt


def image_ncc(pred: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Zero-mean NCC of each image of pred (B, ...) against target (B or 1, ...) of the same
    trailing shape: sum(p0 t0) / (|p0| |t0| + 1e-8) with the means removed -- the NCC of the
    reference's PoU registration test loss (tests/test_cone_beam_fvr.py:24-27, 1 - NCC) -- as
    (B,) values in pred's dtype, on the HIP kernels ``fvr_ncc_*`` (float32 images, float64 sums
    in a fixed order); differentiable in both arguments."""
    if pred.dim() < 2 or target.dim() != pred.dim() or target.shape[1:] != pred.shape[1:] or \
            target.shape[0] not in (1, pred.shape[0]):
        raise ValueError(f"image_ncc: target {tuple(target.shape)} does not match pred {tuple(pred.shape)}")
    _lib.require_device(pred, "image_ncc")
    B = pred.shape[0]
    p = pred.to(torch.float32).reshape(B, -1).contiguous()
    t = target.to(device=pred.device, dtype=torch.float32).reshape(target.shape[0], -1).contiguous()
    return _ImageNCC.apply(p, t).to(pred.dtype)


def sample_isocenter_parity_offsets(vox_spacing) -> torch.Tensor:
    """{-0.5, 0, 0.5}^3 voxel-centre parity offsets scaled by the spacing (cone_beam_fvr.py:122-137)."""
    sp = torch.as_tensor(vox_spacing, dtype=torch.float32)
    sp = sp.repeat(3) if sp.ndim == 0 else sp.flatten()[:3]
    vals = (-0.5, 0.0, 0.5)
    grid = torch.tensor([[a, b, c] for a in vals for b in vals for c in vals], dtype=torch.float32, device=sp.device)
    return grid * sp.unsqueeze(0)


# ---------------------------------------------------------------------------
# fused autograd operator over the PoU kernels
# ---------------------------------------------------------------------------
def _separable_rays(model, rx: torch.Tensor, ry: torch.Tensor, H: int, W: int):
    """(r_x per detector column (W,), r_y per detector row (H,), both monotone) when the shared
    ray arrays are exactly separable (every row of r_x equal, every column of r_y equal: a
    zero-skew pinhole lattice, ``make_pinhole_lattice``), else None.  Cached on the model for the ray
    tensors last seen (object and version), so the check's host synchronisation happens once
    per lattice."""
    key = (id(rx), rx._version, id(ry), ry._version, H, W)
    c = getattr(model, "_sep_rays", None)
    if c is not None and c[0] == key and c[1]() is rx and c[2]() is ry:
        return c[3]
    res = None
    if rx.dim() == 2 and tuple(rx.shape) == (H, W) and tuple(ry.shape) == (H, W):
        if bool((rx == rx[:1]).all()) and bool((ry == ry[:, :1]).all()):
            dev = model.F_blk_ri.device
            rc, rr = _f32(rx[0], dev).contiguous(), _f32(ry[:, 0], dev).contiguous()

            def monotone(a):
                d = a[1:] - a[:-1]
                return bool((d >= 0).all()) or bool((d <= 0).all())

            res = (rc, rr, monotone(rc) and monotone(rr))
    model._sep_rays = (key, weakref.ref(rx), weakref.ref(ry), res)
    return res


class _PouState:
    """What the fused forward leaves for the adjoint (device tensors and shapes).  The tensors
    go through ctx.save_for_backward (``save`` / ``restore``), so that an input changed in place
    between forward and backward raises autograd's version error instead of giving a wrong
    gradient."""
    __slots__ = ("u", "v", "db", "cu", "cv", "rays", "per_pose", "K", "H", "W", "sep", "imgs")

    def tensors(self):
        return (self.u, self.v, self.db, self.cu, self.cv, *self.rays, self.imgs)

    def drop_tensors(self):
        self.u = self.v = self.db = self.cu = self.cv = self.rays = self.imgs = None
        return self

    def restore(self, saved):
        self.u, self.v, self.db, self.cu, self.cv, rx, ry, ic, self.imgs = saved[:9]
        self.rays = (rx, ry, ic)
        return self


def _pou_forward(model, u, v, db, c_u, c_v, rx, ry, ic, out_h, out_w):
    """Slab images (slice + LDS iFFT) and the warp onto the detector for float32 device
    u, v (N, 3) in (x, y, z) voxel order and d_b, cu, cv (B, N): (out (N, H, W), state)."""
    dev = model.F_blk_ri.device
    B, P, N = model.B, model.P, u.shape[0]
    (rxf, ryf, icf), per_pose, K = _rays(rx, ry, ic, N, dev)
    if K != out_h * out_w or tuple(db.shape) != (B, N):
        raise ValueError("PoU render: inconsistent shapes")
    L = _lib.lib()
    st = _lib.stream_ptr(dev)
    Fp = model.packed_blocks()
    imgs = torch.empty((B, N, P, P), device=dev, dtype=torch.float32)
    _lib.check(L.fvr_pou_slice_ifft(Fp.data_ptr(), _VOL_DTYPES[model.F_blk_ri.dtype], B, P, _lib.ptr(u), _lib.ptr(v),
                                    N, _lib.ptr(imgs), st), "fvr_pou_slice_ifft")
    out = torch.empty((N, out_h, out_w), device=dev, dtype=torch.float32)
    sr = _separable_rays(model, rx, ry, out_h, out_w) if not per_pose else None
    inv_ds, c_idx = 1.0 / model.ds_block, float(P // 2)
    if sr is not None and sr[2] and model.separable_warp:
        _lib.check(L.fvr_pou_warp_fwd_sep(_lib.ptr(imgs), B, N, P, _lib.ptr(db), _lib.ptr(c_u), _lib.ptr(c_v),
                                          _lib.ptr(sr[0]), _lib.ptr(sr[1]), _lib.ptr(icf), out_h, out_w, inv_ds, c_idx,
                                          _lib.ptr(out), st), "fvr_pou_warp_fwd_sep")
    else:
        _lib.check(L.fvr_pou_warp_fwd(_lib.ptr(imgs), B, N, P, _lib.ptr(db), _lib.ptr(c_u), _lib.ptr(c_v),
                                      _lib.ptr(rxf), _lib.ptr(ryf), _lib.ptr(icf), per_pose, K, inv_ds, c_idx,
                                      _lib.ptr(out), st), "fvr_pou_warp_fwd")
    sep = None
    if sr is not None and model.separable_adjoint and L.fvr_pou_sep_ok(P, out_h, out_w):
        sep = sr[:2]
    s = _PouState()
    s.u, s.v, s.db, s.cu, s.cv, s.rays, s.per_pose, s.K, s.H, s.W, s.sep, s.imgs = (
        u, v, db, c_u, c_v, (rxf, ryf, icf), per_pose, K, out_h, out_w, sep, imgs)
    return out, s


def _pou_adjoint(model, s, g):
    """The fused adjoint of _pou_forward for the upstream gradient g (N, H, W): per (block,
    pose) the nine sums d/d(d_b, cu, cv), d/du (3), d/dv (3) as part (B, N, 9) and the
    output gradient go (N, K)."""
    dev = s.imgs.device
    B, P, N, K = model.B, model.P, s.u.shape[0], s.K
    L = _lib.lib()
    st = _lib.stream_ptr(dev)
    go = _f32(g, dev).reshape(N, K)
    part = torch.empty((B, N, 9), device=dev, dtype=torch.float32)
    Fp = model.packed_blocks()
    dt = _VOL_DTYPES[model.F_blk_ri.dtype]
    inv_ds, c_idx = 1.0 / model.ds_block, float(P // 2)
    rxf, ryf, icf = s.rays
    if s.sep is not None:
        X = torch.empty((N, K), device=dev, dtype=torch.float32)
        flag = torch.empty(N, device=dev, dtype=torch.float32)
        _lib.check(L.fvr_pou_gprep(_lib.ptr(go), _lib.ptr(icf), N, K, s.per_pose, _lib.ptr(X), _lib.ptr(flag), st),
                   "fvr_pou_gprep")
        rx_col, ry_row = s.sep
        _lib.check(L.fvr_pou_render_bwd_sep(Fp.data_ptr(), dt, B, P, _lib.ptr(s.u), _lib.ptr(s.v), N, _lib.ptr(s.imgs),
                                            _lib.ptr(X), _lib.ptr(flag), _lib.ptr(s.db), _lib.ptr(s.cu),
                                            _lib.ptr(s.cv), _lib.ptr(rx_col), _lib.ptr(ry_row), s.H, s.W, inv_ds,
                                            c_idx, int(model.sep_adjoint_order), _lib.ptr(part), st),
                   "fvr_pou_render_bwd_sep")
    else:
        scale = torch.empty(N, device=dev, dtype=torch.float32)
        _lib.check(L.fvr_pou_gscale(_lib.ptr(go), _lib.ptr(icf), N, K, s.per_pose, _lib.ptr(scale), st),
                   "fvr_pou_gscale")
        _lib.check(L.fvr_pou_render_bwd(Fp.data_ptr(), dt, B, P, _lib.ptr(s.u), _lib.ptr(s.v), N, _lib.ptr(s.imgs),
                                        _lib.ptr(go), _lib.ptr(s.db), _lib.ptr(s.cu), _lib.ptr(s.cv), _lib.ptr(rxf),
                                        _lib.ptr(ryf), _lib.ptr(icf), s.per_pose, K, inv_ds, c_idx,
                                        _lib.ptr(scale), _lib.ptr(part), st), "fvr_pou_render_bwd")
    return part, go


class _PouRender(torch.autograd.Function):
    """Fused slice -> LDS iFFT -> warp for every (block, pose); backward in one fused launch
    (the gather form of the warp adjoint for separable rays, the fixed-point scan otherwise)."""

    @staticmethod
    def forward(ctx, u_xyz, v_xyz, d_b, cu, cv, rx, ry, ic, model, out_h, out_w):
        dev = model.F_blk_ri.device
        out, s = _pou_forward(model, _f32(u_xyz, dev), _f32(v_xyz, dev), _f32(d_b, dev), _f32(cu, dev),
                              _f32(cv, dev), rx, ry, ic, out_h, out_w)
        ctx.save_for_backward(*s.tensors())
        ctx.state, ctx.model = s.drop_tensors(), model
        ctx.meta = (rx.shape, u_xyz.dtype)
        ctx.dts = tuple(t.dtype for t in (d_b, cu, cv, rx, ry, ic))
        return out

    @staticmethod
    def backward(ctx, g):
        s, model = ctx.state.restore(ctx.saved_tensors), ctx.model
        rshape, udt = ctx.meta
        B, P, N, K = model.B, model.P, s.u.shape[0], s.K
        dev = s.imgs.device
        L = _lib.lib()
        st = _lib.stream_ptr(dev)
        part, go = _pou_adjoint(model, s, g)
        guv = torch.empty((N, 6), device=dev, dtype=torch.float32)
        _lib.check(L.fvr_pou_sum_blocks(_lib.ptr(part), B, N, 9, 3, 6, _lib.ptr(guv), st), "fvr_pou_sum_blocks")
        grx = gry = gic = None
        if s.per_pose and any(ctx.needs_input_grad[5:8]):
            rxf, ryf, icf = s.rays
            grx, gry, gic = (torch.empty((N, K), device=dev, dtype=torch.float32) for _ in range(3))
            _lib.check(L.fvr_pou_ray_grads(_lib.ptr(s.imgs), B, N, P, _lib.ptr(go), _lib.ptr(s.db), _lib.ptr(s.cu),
                                           _lib.ptr(s.cv), _lib.ptr(rxf), _lib.ptr(ryf), _lib.ptr(icf), s.per_pose, K,
                                           1.0 / model.ds_block, float(P // 2), _lib.ptr(grx), _lib.ptr(gry),
                                           _lib.ptr(gic), st), "fvr_pou_ray_grads")
            grx, gry, gic = (t.reshape(rshape).to(dt) for t, dt in zip((grx, gry, gic), ctx.dts[3:]))
        return (guv[:, 0:3].to(udt), guv[:, 3:6].to(udt), *(part[..., k].to(ctx.dts[k]) for k in range(3)), grx, gry, gic,
                None, None, None)


class _PouPinhole(torch.autograd.Function):
    """forward_pinhole on a lattice with an SE(3)-invariant ray cache, fused: the pose geometry
    (detector axes and block offsets, cone_beam_fvr.py:507-521, 670-688) in one launch, then
    slice + LDS iFFT and the warp; backward: the fused adjoint, then the block sums and the
    geometry's adjoint to the pose matrices in one launch."""

    @staticmethod
    def forward(ctx, pose_mats, model, lat12, rx, ry, ic, out_h, out_w):
        dev = model.F_blk_ri.device
        M = _f32(pose_mats, dev).contiguous()
        N, B = M.shape[0], model.B
        L = _lib.lib()
        st = _lib.stream_ptr(dev)
        u, v = (torch.empty((N, 3), device=dev, dtype=torch.float32) for _ in range(2))
        db, c_u, c_v = (torch.empty((B, N), device=dev, dtype=torch.float32) for _ in range(3))
        _lib.check(L.fvr_pou_pose_geom(_lib.ptr(M), N, _lib.ptr(lat12), _lib.ptr(model.centers), B, _lib.ptr(u),
                                       _lib.ptr(v), _lib.ptr(db), _lib.ptr(c_u), _lib.ptr(c_v), st),
                   "fvr_pou_pose_geom")
        out, s = _pou_forward(model, u, v, db, c_u, c_v, rx, ry, ic, out_h, out_w)
        ctx.save_for_backward(*s.tensors(), M, lat12)
        ctx.state, ctx.model, ctx.dt = s.drop_tensors(), model, pose_mats.dtype
        return out

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        s, model, M, lat12 = ctx.state.restore(saved), ctx.model, saved[9], saved[10]
        part, _ = _pou_adjoint(model, s, g)
        gM = torch.empty_like(M)
        _lib.check(_lib.lib().fvr_pou_pose_geom_bwd(_lib.ptr(M), M.shape[0], _lib.ptr(lat12),
                                                    _lib.ptr(model.centers), model.B, _lib.ptr(part), _lib.ptr(gM),
                                                    _lib.stream_ptr(M.device)), "fvr_pou_pose_geom_bwd")
        return gM.to(ctx.dt), None, None, None, None, None, None, None


# ---------------------------------------------------------------------------
# the renderer
# ---------------------------------------------------------------------------
class PerspectiveSlabFVR(nn.Module):
    """Differentiable partition-of-unity cone-beam Fourier volume renderer (cone_beam_fvr.py:140-1015)."""

    SLAB_PRESETS = SLAB_PRESETS

    @classmethod
    def resolve_slab_preset(cls, preset, grid_blocks=None, block_res=None, dtype=None):
        if preset is not None:
            if preset not in cls.SLAB_PRESETS:
                raise ValueError(f"Unknown slab preset {preset!r}; expected one of {tuple(cls.SLAB_PRESETS)}")
            cfg = cls.SLAB_PRESETS[preset]
            return (tuple(grid_blocks if grid_blocks is not None else cfg["grid_blocks"]),
                    int(block_res if block_res is not None else cfg["block_res"]),
                    dtype if dtype is not None else cfg["dtype"])
        return (tuple(grid_blocks) if grid_blocks is not None else (2, 2, 2),
                int(block_res if block_res is not None else 64), dtype if dtype is not None else torch.float32)

    def __init__(self, vol_3d, vox_mm: Union[float, Tuple[float, float, float]] = 1.0,
                 origin_world_mm=(0.0, 0.0, 0.0), axis_signs=(1.0, 1.0, 1.0), grid_blocks=(2, 2, 2),
                 block_res: int = 80, pad_mult: int = 2, overlap: float = 0.25, energy_thresh: float = 1e-4,
                 dtype: torch.dtype = torch.float32, use_triton: bool = True, device=None):
        super().__init__()
        if dtype not in _VOL_DTYPES:
            raise TypeError(f"dtype must be float32, float16 or bfloat16, got {dtype}")
        vol = torch.as_tensor(vol_3d).detach()
        if vol.ndim != 3:
            raise ValueError(f"Expected 3D volume tensor, got shape {tuple(vol.shape)}")
        dev = torch.device(device) if device is not None else (vol.device if vol.is_cuda else _lib.default_device())
        if dev.type != "cuda":
            raise RuntimeError("PerspectiveSlabFVR runs on a ROCm device only (there is no CPU path)")
        if isinstance(vox_mm, (tuple, list)):
            sp = tuple(float(x) for x in vox_mm)
            if not (abs(sp[0] - sp[1]) < 1e-3 and abs(sp[1] - sp[2]) < 1e-3):
                raise ValueError(f"Volume spacing must be isotropic (got {sp})")
            self.vox_mm = sp[0]
        else:
            self.vox_mm = float(vox_mm)
        self.use_triton = bool(use_triton)
        self.block_res = int(block_res)
        self.pad_mult = int(pad_mult)
        self.P = self.block_res * self.pad_mult
        vol = vol.to(device=dev, dtype=torch.float64)
        n = vol.shape
        wl = [_make_pou_1d(int(nk), int(k), overlap, dev) for nk, k in zip(n, grid_blocks)]
        ov = [max(2, int(round(s * overlap))) if k > 1 else 0 for (_, s), k in zip(wl, grid_blocks)]
        m = [int(np.ceil(s + 2 * o)) for (_, s), o in zip(wl, ov)]
        M = max(m) + (1 if max(grid_blocks) > 1 else 0)
        self.ds_block = (M * self.vox_mm) / float(self.block_res)
        ws = _sinc2_precomp_1d(self.block_res, self.P, dev)
        w3 = ws[:, None, None] * ws[None, :, None] * ws[None, None, :]
        pad_b = (self.P - self.block_res) // 2
        half_vox_mm = ((self.P // 2) - (pad_b + (self.block_res - 1) * 0.5)) * self.ds_block

        def support(w):
            nz = torch.nonzero(w > 1e-6).squeeze(-1)
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
                    l0, l1, l2 = sub.shape
                    q0, q1, q2 = (M - l0) // 2, (M - l1) // 2, (M - l2) // 2
                    cube = F.pad(sub, (q2, M - l2 - q2, q1, M - l1 - q1, q0, M - l0 - q0))
                    cidx = (a0 + (M - 1) * 0.5 - q0, a1 + (M - 1) * 0.5 - q1, a2 + (M - 1) * 0.5 - q2)
                    centers.append([origin_world_mm[k] + axis_signs[k] * cidx[k] * self.vox_mm
                                    + axis_signs[k] * half_vox_mm for k in range(3)])
                    if M != self.block_res:
                        cube = F.interpolate(cube[None, None], size=(self.block_res,) * 3, mode="trilinear",
                                             align_corners=False)[0, 0]
                    pb = (pad_b, self.P - self.block_res - pad_b)
                    cubes.append(F.pad(cube * self.ds_block * w3, pb * 3))
        if not cubes:
            raise ValueError("Input volume contains no non-zero attenuation mass above threshold.")
        self.B = len(cubes)
        Fb = torch.empty((self.B, 2, self.P, self.P, self.P), device=dev, dtype=dtype)
        for b, c in enumerate(cubes):  # one block at a time keeps the float64 FFT's footprint small
            spec = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(c)))
            Fb[b, 0] = spec.real.to(dtype)
            Fb[b, 1] = spec.imag.to(dtype)
            del spec
        del cubes
        self.register_buffer("F_blk_ri", Fb)
        self.register_buffer("centers", torch.tensor(centers, dtype=torch.float32, device=dev))
        freqs = torch.fft.fftshift(torch.fft.fftfreq(self.P, d=self.ds_block)).to(dev)
        fy, fx = torch.meshgrid(freqs, freqs, indexing="ij")
        self.register_buffer("FX", fx.float())
        self.register_buffer("FY", fy.float())
        self.register_buffer("FX_flat", fx.reshape(-1).float().contiguous())
        self.register_buffer("FY_flat", fy.reshape(-1).float().contiguous())

    @classmethod
    def from_volume(cls, vol_zyx, vox_mm: float = 1.0, grid_blocks=None, block_res=None, pad_mult: int = 2,
                    overlap: float = 0.25, center_mode: str = "pixel_center", preset: Optional[str] = None,
                    dtype: Optional[torch.dtype] = None, **kwargs) -> "PerspectiveSlabFVR":
        """From an isotropic (Z, Y, X) volume centred at the world origin (cone_beam_fvr.py:332-374)."""
        gb, br, dt = cls.resolve_slab_preset(preset, grid_blocks, block_res, dtype)
        vt = torch.as_tensor(vol_zyx, dtype=torch.float32)
        nz, ny, nx = vt.shape
        v = float(vox_mm)
        if center_mode == "pixel_center":
            origin = (-0.5 * (nz - 1) * v, -0.5 * (ny - 1) * v, -0.5 * (nx - 1) * v)
        elif center_mode == "grid_center":
            origin = (-0.5 * nz * v, -0.5 * ny * v, -0.5 * nx * v)
        else:
            raise ValueError(f"Unknown center_mode={center_mode!r}; expected 'pixel_center' or 'grid_center'")
        return cls(vt, vox_mm=v, origin_world_mm=origin, axis_signs=(1.0, 1.0, 1.0), grid_blocks=gb, block_res=br,
                   pad_mult=pad_mult, overlap=overlap, dtype=dt, **kwargs)

    def to_precision(self, dtype: torch.dtype) -> "PerspectiveSlabFVR":
        """Store the block spectra in ``dtype`` (float32 / float16 / bfloat16; cone_beam_fvr.py:305-314)."""
        if dtype not in _VOL_DTYPES:
            raise TypeError(f"dtype must be float32, float16 or bfloat16, got {dtype}")
        self.F_blk_ri = self.F_blk_ri.to(dtype=dtype)
        return self

    def compile_renderer(self, mode: str = "reduce-overhead", fullgraph: bool = False) -> "PerspectiveSlabFVR":
        """Accepted for call compatibility (cone_beam_fvr.py:316-330): the path is already fused HIP launches."""
        return self

    # -- building blocks -----------------------------------------------------------
    # the fused LDS kernels are used whenever P allows them; False selects the
    # generic composition (slice operator + torch.fft + warp), for cross-checks
    fused_kernels = True
    # separable shared rays take the gather form of the fused adjoint; False selects the
    # fixed-point scan for them too, for cross-checks
    separable_adjoint = True
    # workgroup order of the gather form: 0 block-major, 1 pose-major
    sep_adjoint_order = 0
    # separable monotone shared rays take the tiled forward warp (the same projections bit for
    # bit); False selects the per-pixel warp, for cross-checks
    separable_warp = True

    def fused_ok(self) -> bool:
        return bool(self.fused_kernels) and bool(_lib.lib().fvr_pou_lds_ok(self.P))

    def packed_blocks(self) -> torch.Tensor:
        """The block spectra in the fused kernels' gather layout (interleaved (re, im), zero
        margins: ``fvr_pou_pack_blocks``), cached and rebuilt whenever ``F_blk_ri`` changes."""
        Fb = self.F_blk_ri
        key = (Fb.data_ptr(), Fb.dtype, Fb._version, tuple(Fb.shape))
        if getattr(self, "_pk_key", None) != key:
            L = _lib.lib()
            dt = _VOL_DTYPES[Fb.dtype]
            Fc = Fb.contiguous()
            buf = torch.empty(int(L.fvr_pou_packed_bytes(dt, self.B, self.P)), dtype=torch.uint8, device=Fb.device)
            _lib.check(L.fvr_pou_pack_blocks(Fc.data_ptr(), dt, self.B, self.P, buf.data_ptr(),
                                             _lib.stream_ptr(Fb.device)), "fvr_pou_pack_blocks")
            self._pk, self._pk_key = buf, key
        return self._pk

    def _slice_and_ifft(self, u_det: torch.Tensor, v_det: torch.Tensor) -> torch.Tensor:
        """(B, N, P, P) real slab images (cone_beam_fvr.py:480-505)."""
        dev = self.F_blk_ri.device
        u, v = u_det.to(dev).flip(-1), v_det.to(dev).flip(-1)
        N, B, P = u.shape[0], self.B, self.P
        if self.fused_ok() and not (u.requires_grad or v.requires_grad):
            imgs = torch.empty((B, N, P, P), device=dev, dtype=torch.float32)
            _lib.check(_lib.lib().fvr_pou_slice_ifft(self.packed_blocks().data_ptr(), _VOL_DTYPES[self.F_blk_ri.dtype],
                                                     B, P, _lib.ptr(_f32(u, dev)), _lib.ptr(_f32(v, dev)), N,
                                                     _lib.ptr(imgs), _lib.stream_ptr(dev)), "fvr_pou_slice_ifft")
            return imgs
        scale = float(P) * self.ds_block
        s_ri = fourier_slice_3d(self.F_blk_ri, u * scale, v * scale, self.FX_flat, self.FY_flat,
                                (float(P // 2),) * 3)
        s_ri = s_ri.float().view(B, N, 2, P, P)
        sl = torch.complex(s_ri[:, :, 0], s_ri[:, :, 1])
        return torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(sl, dim=(-2, -1)), dim=(-2, -1)),
                                  dim=(-2, -1)).real

    def _compute_block_offsets(self, srcs, u_det, v_det, n_0):
        """(d_b, cu, cv) each (B, N) (cone_beam_fvr.py:507-521)."""
        diff = self.centers.view(self.B, 1, 3) - srcs.view(1, -1, 3)
        return ((diff * n_0.view(1, -1, 3)).sum(-1), (diff * u_det.view(1, -1, 3)).sum(-1),
                (diff * v_det.view(1, -1, 3)).sum(-1))

    @staticmethod
    def _make_ray_cache(src0, trg0, u0, v0, n0):
        """SE(3)-invariant (rx, ry, inv_cos) (H, W) and their flat copies (cone_beam_fvr.py:524-545)."""
        rays = trg0 - src0.view(1, 1, 3)
        dt = rays.dtype if rays.dtype == torch.float64 else torch.float32
        rays = rays.to(dt)
        r = rays * torch.rsqrt((rays * rays).sum(dim=-1, keepdim=True).clamp(min=1e-12))
        ic = (1.0 / (r * n0.to(dt).view(1, 1, 3)).sum(dim=-1).clamp(min=0.1)).contiguous()
        rx = ((r * u0.to(dt).view(1, 1, 3)).sum(dim=-1) * ic).contiguous()
        ry = ((r * v0.to(dt).view(1, 1, 3)).sum(dim=-1) * ic).contiguous()
        return rx, ry, ic, rx.reshape(-1).contiguous(), ry.reshape(-1).contiguous(), ic.reshape(-1).contiguous()

    @staticmethod
    def _extract_detector_basis(srcs, trg_pts, u_det=None, v_det=None, n_0=None):
        """Orthonormal detector axes (N, 3) each (cone_beam_fvr.py:576-596)."""
        c_det = trg_pts.mean(dim=(1, 2))
        if u_det is None:
            ur = trg_pts[:, 0, -1] - trg_pts[:, 0, 0]
            u_det = ur / (ur.norm(dim=-1, keepdim=True) + 1e-12)
        if v_det is None:
            vr = trg_pts[:, -1, 0] - trg_pts[:, 0, 0]
            v_det = vr / (vr.norm(dim=-1, keepdim=True) + 1e-12)
        if n_0 is None:
            nr = torch.cross(v_det, u_det, dim=-1)
            n_0 = nr / (nr.norm(dim=-1, keepdim=True) + 1e-12)
            flip = ((c_det - srcs) * n_0).sum(dim=-1, keepdim=True) < 0
            n_0 = torch.where(flip, -n_0, n_0)
        return u_det, v_det, n_0

    def _render(self, srcs, u_det, v_det, n_0, rx, ry, ic, out_h, out_w) -> torch.Tensor:
        """(N, H, W): fused path for LDS-sized P, generic composition otherwise."""
        dev = self.F_blk_ri.device
        srcs, u_det, v_det, n_0 = (t.to(dev) for t in (srcs, u_det, v_det, n_0))
        d_b, cu, cv = self._compute_block_offsets(srcs, u_det, v_det, n_0)
        if self.fused_ok():
            return _PouRender.apply(u_det.flip(-1), v_det.flip(-1), d_b, cu, cv, rx.to(dev), ry.to(dev),
                                    ic.to(dev), self, out_h, out_w)
        imgs = self._slice_and_ifft(u_det, v_det)
        return _PouWarpSum.apply(imgs, d_b, cu, cv, rx.to(dev), ry.to(dev), ic.to(dev), 1.0 / self.ds_block,
                                 float(self.P // 2), out_h, out_w)

    def _project_core_invariant(self, srcs, u_det, v_det, n_0, ray_cache) -> torch.Tensor:
        """Projection with SE(3)-invariant camera-frame rays (cone_beam_fvr.py:547-573)."""
        rx_hw, ry_hw, ic_hw = ray_cache[:3]
        return self._render(srcs, u_det, v_det, n_0, rx_hw, ry_hw, ic_hw, int(rx_hw.shape[0]), int(rx_hw.shape[1]))

    def _project_core(self, srcs, trg_pts, u_det, v_det, n_0) -> torch.Tensor:
        """Projection with explicit per-pose rays (cone_beam_fvr.py:423-478)."""
        N, H, W, _ = trg_pts.shape
        rays = trg_pts - srcs.view(N, 1, 1, 3)
        r = rays * torch.rsqrt((rays * rays).sum(dim=-1, keepdim=True).clamp(min=1e-12))
        ic = 1.0 / (r * n_0.view(N, 1, 1, 3)).sum(dim=-1).clamp(min=0.1)
        rx = (r * u_det.view(N, 1, 1, 3)).sum(dim=-1) * ic
        ry = (r * v_det.view(N, 1, 1, 3)).sum(dim=-1) * ic
        return self._render(srcs, u_det, v_det, n_0, rx, ry, ic, H, W)

    # -- public API ------------------------------------------------------------------
    def project_rays_batched(self, srcs, trg_pts, u_det=None, v_det=None, n_0=None) -> torch.Tensor:
        """(N, H, W) cone-beam line integrals for sources (N, 3) and lattices (N, H, W, 3) (:598-616)."""
        u_det, v_det, n_0 = self._extract_detector_basis(srcs, trg_pts, u_det, v_det, n_0)
        return self._project_core(srcs, trg_pts, u_det, v_det, n_0)

    def project_rays(self, src, trg_pts, u_det=None, v_det=None, n_0=None) -> torch.Tensor:
        """(H, W) line integrals for one source (3,) and lattice (H, W, 3) (:618-636)."""
        return self.project_rays_batched(src.unsqueeze(0), trg_pts.unsqueeze(0),
                                         None if u_det is None else u_det.unsqueeze(0),
                                         None if v_det is None else v_det.unsqueeze(0),
                                         None if n_0 is None else n_0.unsqueeze(0)).squeeze(0)

    @staticmethod
    def make_pinhole_lattice(k_inv: torch.Tensor, sdd_mm: float, height: int, width: int):
        """Camera-frame pinhole lattice + SE(3)-invariant ray cache (11-tuple) (:706-728).

        Formed in float64 and returned in float32.  For a zero-skew pinhole r_x depends on
        the detector column only and r_y on the row only; when the float64 cache shows that
        (to 1e-12 relative), every row of r_x is made row 0's and every column of r_y
        column 0's, so that the float32 cache is exactly separable and the backward takes
        the gather form of the warp adjoint (``fvr_pou_render_bwd_sep``)."""
        k0 = k_inv.squeeze(0) if k_inv.ndim == 3 else k_inv
        dev, dt = k0.device, torch.float64
        cols = torch.arange(width, device=dev, dtype=dt) + 0.5
        rows = torch.arange(height, device=dev, dtype=dt) + 0.5
        rr, cc = torch.meshgrid(rows, cols, indexing="ij")
        pix = torch.stack([cc, rr, torch.ones_like(cc)], dim=-1)
        dirs = torch.einsum("ij,hwj->hwi", k0.to(dtype=dt), pix)
        trg0 = dirs * (float(sdd_mm) / dirs[..., 2:3].clamp(min=1e-6))
        src0 = torch.zeros(3, device=dev, dtype=dt)
        u0, v0, n0 = PerspectiveSlabFVR._extract_detector_basis(src0.unsqueeze(0), trg0.unsqueeze(0))
        u0, v0, n0 = u0.squeeze(0), v0.squeeze(0), n0.squeeze(0)
        rx, ry, ic = PerspectiveSlabFVR._make_ray_cache(src0, trg0, u0, v0, n0)[:3]
        tol = 1e-12 * max(1.0, float(rx.abs().max()), float(ry.abs().max()))
        if float((rx - rx[:1]).abs().max()) <= tol and float((ry - ry[:, :1]).abs().max()) <= tol:
            rx, ry = rx[:1].expand_as(rx), ry[:, :1].expand_as(ry)
        f32 = torch.float32
        rx, ry, ic = (t.to(f32).contiguous() for t in (rx, ry, ic))
        return (src0.to(f32), trg0.to(f32), u0.to(f32), v0.to(f32), n0.to(f32), rx, ry, ic,
                rx.reshape(-1).contiguous(), ry.reshape(-1).contiguous(), ic.reshape(-1).contiguous())

    def _lattice_frame(self, lattice) -> torch.Tensor:
        """(src0, u0, v0, n0) of a lattice as one float32 device array of 12, cached for the
        lattice tensors last seen (object and version)."""
        ts = lattice[0], lattice[2], lattice[3], lattice[4]
        key = tuple((id(t), t._version) for t in ts)
        c = getattr(self, "_lat12", None)
        if c is not None and c[0] == key and all(r() is t for r, t in zip(c[1], ts)):
            return c[2]
        lat12 = torch.cat([_f32(t, self.F_blk_ri.device).reshape(3) for t in ts]).contiguous()
        self._lat12 = (key, [weakref.ref(t) for t in ts], lat12)
        return lat12

    def _forward_from_lattice(self, pose_mats: torch.Tensor, lattice) -> torch.Tensor:
        """Transform a camera-frame lattice by SE(3) matrices (N, 4, 4) and project (:670-688)."""
        dev = self.F_blk_ri.device
        pose_mats = pose_mats.to(dev)
        if len(lattice) >= 11 and self.fused_ok() and pose_mats.dim() == 3 and tuple(pose_mats.shape[1:]) == (4, 4):
            rx, ry, ic = (x.to(dev) for x in lattice[5:8])
            return _PouPinhole.apply(pose_mats, self, self._lattice_frame(lattice), rx, ry, ic, int(rx.shape[-2]),
                                     int(rx.shape[-1])).unsqueeze(1)
        src0, trg0, u0, v0, n0 = (t.to(dev) for t in lattice[:5])
        R, t = pose_mats[:, :3, :3], pose_mats[:, :3, 3]
        srcs = torch.einsum("nij,j->ni", R, src0) + t
        u_det = torch.einsum("nij,j->ni", R, u0)
        v_det = torch.einsum("nij,j->ni", R, v0)
        n_0 = torch.einsum("nij,j->ni", R, n0)
        if len(lattice) >= 11:
            return self._project_core_invariant(srcs, u_det, v_det, n_0, [x.to(dev) for x in lattice[5:11]]).unsqueeze(1)
        trg = torch.einsum("nij,hwj->nhwi", R, trg0) + t[:, None, None, :]
        return self.project_rays_batched(srcs, trg, u_det=u_det, v_det=v_det, n_0=n_0).unsqueeze(1)

    def forward_pinhole(self, rt_inv: torch.Tensor, lattice) -> torch.Tensor:
        """(N, 1, H, W) pinhole projections for SE(3) matrices ``rt_inv`` (N, 4, 4) (:730-740)."""
        return self._forward_from_lattice(rt_inv, lattice)

    def warmup_pinhole(self, lattice, batch_size: int = 1, backward: bool = True) -> None:
        """Run one forward (and backward) so later calls do no first-use set-up (:742-759)."""
        dev = self.F_blk_ri.device
        eye = torch.eye(4, device=dev, dtype=torch.float32).unsqueeze(0).repeat(max(1, int(batch_size)), 1, 1)
        if backward:
            eye.requires_grad_(True)
            self.forward_pinhole(eye, lattice).sum().backward()
        else:
            with torch.no_grad():
                self.forward_pinhole(eye, lattice)
        torch.cuda.synchronize(dev)

    def forward_carm_view(self, yaw_deg, pitch_deg, roll_deg, tx_vox, ty_vox, tz_vox, sad_mm: float, sdd_mm: float,
                          det_px: int, det_fov_mm: float, view: str = "ap") -> torch.Tensor:
        """(det_px, det_px) C-arm projection, AP or lateral geometry (cone_beam_fvr.py:761-832).

        The geometry (rotation, source, lattice, rays) is formed in float64 and cast to
        float32 only where the kernels read it: a translation moves the source and every
        lattice point together, so its gradient is the difference of two large sums over
        the pixels' ray gradients, which float32 autograd would round to ~1e-3."""
        dev = self.F_blk_ri.device
        f64 = torch.float64
        a = [torch.as_tensor(x).to(device=dev, dtype=f64) for x in (yaw_deg, pitch_deg, roll_deg, tx_vox, ty_vox,
                                                                    tz_vox)]
        yr, pr, rr = (torch.deg2rad(x) for x in a[:3])
        cy, sy, cp, sp, cr, sr = torch.cos(yr), torch.sin(yr), torch.cos(pr), torch.sin(pr), torch.cos(rr), torch.sin(rr)
        z, o = torch.zeros((), device=dev, dtype=f64), torch.ones((), device=dev, dtype=f64)
        rz = torch.stack([torch.stack([o, z, z]), torch.stack([z, cy, -sy]), torch.stack([z, sy, cy])])
        ry = torch.stack([torch.stack([cr, z, sr]), torch.stack([z, o, z]), torch.stack([-sr, z, cr])])
        rx = torch.stack([torch.stack([cp, -sp, z]), torch.stack([sp, cp, z]), torch.stack([z, z, o])])
        rot = rz @ ry @ rx
        lin = torch.linspace(-0.5, 0.5, det_px, device=dev, dtype=f64) * float(det_fov_mm)
        vv, uu = torch.meshgrid(lin, lin, indexing="ij")
        det_d = float(sdd_mm - sad_mm)
        if view == "ap":
            src0 = torch.tensor([0.0, -float(sad_mm), 0.0], device=dev, dtype=f64)
            trg0 = torch.stack([vv, torch.full_like(uu, det_d), uu], dim=-1)
            u0, v0, n0 = (torch.tensor(x, device=dev, dtype=f64) for x in ([0.0, 0.0, 1.0], [1.0, 0.0, 0.0],
                                                                          [0.0, 1.0, 0.0]))
        else:
            src0 = torch.tensor([0.0, 0.0, -float(sad_mm)], device=dev, dtype=f64)
            trg0 = torch.stack([vv, uu, torch.full_like(uu, det_d)], dim=-1)
            u0, v0, n0 = (torch.tensor(x, device=dev, dtype=f64) for x in ([0.0, 1.0, 0.0], [1.0, 0.0, 0.0],
                                                                          [0.0, 0.0, 1.0]))
        shift = torch.stack([a[5], a[4], a[3]]) * self.vox_mm
        src = rot @ src0 - shift
        trg = torch.einsum("ij,hwj->hwi", rot, trg0) - shift
        return self.project_rays(src, trg, u_det=rot @ u0, v_det=rot @ v0, n_0=rot @ n0)
