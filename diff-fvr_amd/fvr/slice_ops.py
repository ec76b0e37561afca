"""The reference's generic slice operator on its planar layout, backed by HIP.

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
ship).
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
