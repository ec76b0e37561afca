"""The reference's operator module (``src/fvr/triton_slice.py``), backed by HIP.

``fourier_slice_3d(vol_ri, u_xyz, v_xyz, fx, fy, center_xyz)`` restates
``triton_slice.py:648-661``: ``vol_ri`` is ``(B_vol, 2, D, H, W)`` (real, imag
planes) in float32 / float16 / bfloat16, sample ``q`` of pose ``n`` sits at
``center + fx[q] u[n] + fy[q] v[n]`` (x = W, y = H, z = D voxel axes), and is
interpolated trilinearly with every out-of-range corner read as zero
(``grid_sample(align_corners=True, padding_mode='zeros')``).  The output is
``(B_vol, N, 2, K)`` in the volume's dtype.  Autograd returns gradients for
``u_xyz`` and ``v_xyz`` only, like the reference's kernel path
(``triton_slice.py:306-339``); a volume that requires grad is refused (the
reference falls back to ``grid_sample`` there, which this package does not
ship).  Kernels: ``fvr_slice_fwd`` / ``fvr_slice_bwd`` (csrc/fvr_slice.hip).

``pou_pinhole_warp_sum(imgs, d_b, cu, cv, rx, ry, inv_cos, ...)`` restates
``triton_slice.py:787-808``: the (N, H, W) sum over blocks of the bilinear sample
of each slab image at the detector pixel's position, times 1 / cos, with
gradients for the images, d_b, cu, cv and per-pose rays.  Kernels:
``fvr_pou_warp_fwd`` / ``fvr_pou_warp_bwd`` / ``fvr_pou_ray_grads``
(csrc/fvr_pou.hip; the image gradient accumulated in 64-bit fixed point, so it
is bit-identical run to run).
"""

from __future__ import annotations

from typing import Tuple

import torch

from . import _lib

_VOL_DTYPES = {torch.float32: 0, torch.float16: 1, torch.bfloat16: 2}


def _check(vol_ri, u_xyz, v_xyz, fx, fy):
    _lib.require_device(vol_ri, "fourier_slice_3d")
    if vol_ri.dim() != 5 or vol_ri.shape[1] != 2:
        raise ValueError(f"vol_ri must be (B_vol, 2, D, H, W), got {tuple(vol_ri.shape)}")
    if vol_ri.dtype not in _VOL_DTYPES:
        raise TypeError(f"vol_ri dtype {vol_ri.dtype} not supported (float32, float16, bfloat16)")
    if u_xyz.dim() != 2 or u_xyz.shape[1] != 3 or tuple(v_xyz.shape) != tuple(u_xyz.shape):
        raise ValueError("u_xyz and v_xyz must both be (N, 3)")
    if fx.numel() != fy.numel() or fx.numel() == 0:
        raise ValueError("fx and fy must be non-empty and of equal length")


class _FourierSlice3D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol_ri, u_xyz, v_xyz, fx, fy, cx, cy, cz):
        dev = vol_ri.device
        vol = vol_ri.contiguous()
        u = u_xyz.detach().to(device=dev, dtype=torch.float32).contiguous()
        v = v_xyz.detach().to(device=dev, dtype=torch.float32).contiguous()
        fxf = fx.detach().to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
        fyf = fy.detach().to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
        B, _, D, H, W = vol.shape
        N, K = u.shape[0], fxf.numel()
        out = torch.empty((B, N, 2, K), device=dev, dtype=vol.dtype)
        _lib.check(_lib.lib().fvr_slice_fwd(vol.data_ptr(), _VOL_DTYPES[vol.dtype], B, D, H, W, _lib.ptr(u),
                                            _lib.ptr(v), N, _lib.ptr(fxf), _lib.ptr(fyf), K, float(cx), float(cy),
                                            float(cz), out.data_ptr(), _lib.stream_ptr(dev)), "fvr_slice_fwd")
        ctx.save_for_backward(vol, u, v, fxf, fyf)
        ctx.center = (float(cx), float(cy), float(cz))
        ctx.uv_dtypes = (u_xyz.dtype, v_xyz.dtype, u_xyz.device)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        vol, u, v, fxf, fyf = ctx.saved_tensors
        cx, cy, cz = ctx.center
        dev = vol.device
        B, _, D, H, W = vol.shape
        N, K = u.shape[0], fxf.numel()
        go = grad_out.to(dtype=vol.dtype).contiguous()
        L = _lib.lib()
        part = torch.empty(int(L.fvr_slice_bwd_workspace(N, K)), device=dev, dtype=torch.float32)
        gu = torch.empty((N, 3), device=dev, dtype=torch.float32)
        gv = torch.empty((N, 3), device=dev, dtype=torch.float32)
        _lib.check(L.fvr_slice_bwd(vol.data_ptr(), _VOL_DTYPES[vol.dtype], B, D, H, W, go.data_ptr(), _lib.ptr(u),
                                   _lib.ptr(v), N, _lib.ptr(fxf), _lib.ptr(fyf), K, cx, cy, cz, _lib.ptr(part),
                                   _lib.ptr(gu), _lib.ptr(gv), _lib.stream_ptr(dev)), "fvr_slice_bwd")
        ud, vd, udev = ctx.uv_dtypes
        return None, gu.to(device=udev, dtype=ud), gv.to(device=udev, dtype=vd), None, None, None, None, None


def fourier_slice_3d(vol_ri: torch.Tensor, u_xyz: torch.Tensor, v_xyz: torch.Tensor, fx: torch.Tensor,
                     fy: torch.Tensor, center_xyz: Tuple[float, float, float], use_triton: bool = True) -> torch.Tensor:
    """(B_vol, N, 2, K) slices of ``vol_ri`` (see the module docstring).

    ``use_triton`` is accepted for call compatibility and ignored: the HIP
    kernel is the only implementation.
    """
    _check(vol_ri, u_xyz, v_xyz, fx, fy)
    if vol_ri.requires_grad:
        raise NotImplementedError("fourier_slice_3d: gradients w.r.t. the spectrum are not provided "
                                  "(u_xyz / v_xyz only, as in the reference's kernel path)")
    cx, cy, cz = center_xyz
    return _FourierSlice3D.apply(vol_ri, u_xyz, v_xyz, fx, fy, float(cx), float(cy), float(cz))


def triton_slice_available(vol_ri: torch.Tensor | None = None) -> bool:
    """True when the HIP library is loadable and ``vol_ri`` (if given) is on a ROCm device."""
    try:
        _lib.lib()
    except ImportError:
        return False
    return vol_ri is None or bool(vol_ri.is_cuda)


# ---------------------------------------------------------------------------
# pou_pinhole_warp_sum: autograd over the warp kernels
# ---------------------------------------------------------------------------
def _f32(t: torch.Tensor, dev) -> torch.Tensor:
    return t.detach().to(device=dev, dtype=torch.float32).contiguous()


def _rays(rx, ry, ic, N: int, dev):
    """Flattened float32 ray arrays and the per-pose flag: (K,) shared or (N, K) per pose."""
    per_pose = rx.dim() == 3
    r = [_f32(a, dev) for a in (rx, ry, ic)]
    if per_pose and r[0].shape[0] != N:
        raise ValueError(f"per-pose ray arrays hold {r[0].shape[0]} poses, expected {N}")
    if per_pose:
        K = r[0].shape[-2] * r[0].shape[-1]
        r = [a.reshape(N, K) for a in r]
    else:
        K = r[0].numel()
        r = [a.reshape(K) for a in r]
    return r, int(per_pose), K


class _PouWarpSum(torch.autograd.Function):
    """pou_pinhole_warp_sum (triton_slice.py:787-808): grads for imgs, d_b, cu, cv and (per-pose) rays."""

    @staticmethod
    def forward(ctx, imgs, d_b, cu, cv, rx, ry, ic, inv_ds, c_idx, out_h, out_w):
        dev = imgs.device
        B, N, P, P2 = imgs.shape
        if P2 != P:
            raise ValueError("slab images must be square (B, N, P, P)")
        im = _f32(imgs, dev)
        db, u_, v_ = _f32(d_b, dev), _f32(cu, dev), _f32(cv, dev)
        (rxf, ryf, icf), per_pose, K = _rays(rx, ry, ic, N, dev)
        if K != out_h * out_w or tuple(db.shape) != (B, N) or tuple(u_.shape) != (B, N) or tuple(v_.shape) != (B, N):
            raise ValueError("pou_pinhole_warp_sum: inconsistent shapes")
        out = torch.empty((N, out_h, out_w), device=dev, dtype=torch.float32)
        _lib.check(_lib.lib().fvr_pou_warp_fwd(_lib.ptr(im), B, N, P, _lib.ptr(db), _lib.ptr(u_), _lib.ptr(v_),
                                               _lib.ptr(rxf), _lib.ptr(ryf), _lib.ptr(icf), per_pose, K,
                                               float(inv_ds), float(c_idx), _lib.ptr(out), _lib.stream_ptr(dev)),
                   "fvr_pou_warp_fwd")
        ctx.save_for_backward(im, db, u_, v_, rxf, ryf, icf)
        ctx.meta = (B, N, P, K, per_pose, float(inv_ds), float(c_idx), imgs.dtype, (out_h, out_w), rx.shape)
        ctx.dts = tuple(t.dtype for t in (d_b, cu, cv, rx, ry, ic))
        return out

    @staticmethod
    def backward(ctx, g):
        im, db, u_, v_, rxf, ryf, icf = ctx.saved_tensors
        B, N, P, K, per_pose, inv_ds, c_idx, idt, hw, rshape = ctx.meta
        dev = im.device
        L = _lib.lib()
        st = _lib.stream_ptr(dev)
        go = _f32(g, dev).reshape(N, K)
        want_img = ctx.needs_input_grad[0]
        part = torch.empty((B, N, 3), device=dev, dtype=torch.float32)
        scale = torch.empty(N, device=dev, dtype=torch.float32)
        gimg = None
        ws = None
        if want_img:
            gimg = torch.empty((B, N, P, P), device=dev, dtype=torch.float32)
            ws = torch.empty(int(L.fvr_pou_warp_bwd_workspace(B, N, P)), device=dev, dtype=torch.uint8)
            _lib.check(L.fvr_pou_gscale(_lib.ptr(go), _lib.ptr(icf), N, K, per_pose, _lib.ptr(scale), st),
                       "fvr_pou_gscale")
        _lib.check(L.fvr_pou_warp_bwd(_lib.ptr(im), B, N, P, _lib.ptr(go), _lib.ptr(db), _lib.ptr(u_), _lib.ptr(v_),
                                      _lib.ptr(rxf), _lib.ptr(ryf), _lib.ptr(icf), per_pose, K, inv_ds, c_idx,
                                      _lib.ptr(scale), _lib.ptr(ws) if ws is not None else None,
                                      _lib.ptr(gimg) if gimg is not None else None, _lib.ptr(part), st),
                   "fvr_pou_warp_bwd")
        grx = gry = gic = None
        if per_pose and any(ctx.needs_input_grad[4:7]):
            grx, gry, gic = (torch.empty((N, K), device=dev, dtype=torch.float32) for _ in range(3))
            _lib.check(L.fvr_pou_ray_grads(_lib.ptr(im), B, N, P, _lib.ptr(go), _lib.ptr(db), _lib.ptr(u_),
                                           _lib.ptr(v_), _lib.ptr(rxf), _lib.ptr(ryf), _lib.ptr(icf), per_pose, K,
                                           inv_ds, c_idx, _lib.ptr(grx), _lib.ptr(gry), _lib.ptr(gic), st),
                       "fvr_pou_ray_grads")
            grx, gry, gic = (t.reshape(rshape).to(dt) for t, dt in zip((grx, gry, gic), ctx.dts[3:]))
        return (gimg.to(idt) if gimg is not None else None, *(part[..., k].to(ctx.dts[k]) for k in range(3)),
                grx, gry, gic, None, None, None, None)


def pou_pinhole_warp_sum(imgs, d_b, cu, cv, rx_hw, ry_hw, inv_cos_hw, rx_flat=None, ry_flat=None, ic_flat=None,
                         inv_ds: float = 1.0, c_idx: float = 0.0, use_triton: bool = True) -> torch.Tensor:
    """(N, H, W) = sum_b bilinear(imgs[b, n], p_bn) * inv_cos (triton_slice.py:787-808), on HIP.

    Ray arrays are (H, W) shared by all poses (the reference's signature) or
    (N, H, W) per pose; ``use_triton`` is accepted and ignored.
    """
    _lib.require_device(imgs, "pou_pinhole_warp_sum")
    H, W = int(rx_hw.shape[-2]), int(rx_hw.shape[-1])
    return _PouWarpSum.apply(imgs, d_b, cu, cv, rx_hw, ry_hw, inv_cos_hw, float(inv_ds), float(c_idx), H, W)
