"""Drop-in ``TorchDiffFVREngine`` (the reference's torch path, convention T) on HIP.

Restates ``src/fvr/torch_engine.py`` of pareelamre/diff-fvr:

* spectrum: centred zero padding, ``fftshift(fftn(.))`` (no ifftshift, no
  sinc^2), stored as ``f_vol_5d`` = (1, 2, D, H, W) real/imag planes
  (:28-88); computed here in float64 on the device, then cast;
* ``_extract_slice_spectrum`` / ``forward`` (:176-253): the slice through the
  generic HIP operator (``triton_slice.fourier_slice_3d``), the phase ramp, and for
  ``forward`` the inverse FFT + ReLU, differentiable w.r.t. the poses;
* ``forward_spectral_ncc`` (:269-307), the registration hot path: one fused HIP
  launch per call (slice + phase ramp + high-pass + NCC sums + exact pose
  adjoint, ``csrc/fvr_sncc.hip``) behind a ``torch.autograd.Function`` whose
  backward returns d loss / d poses from moments reduced in float64.

Poses may live on the host or the device; results are on the poses' device.
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Tuple, Union

import numpy as np
import torch
import torch.nn as nn

from . import _items, _lib
from .triton_slice import fourier_slice_3d
from .sncc import _Workspace, launch_items, tile_rows
from .spectrum import PackedSpectrum

_STORAGE_OF = {torch.float32: "f32", torch.float16: "fp16", torch.bfloat16: "bf16"}
NCC_EPS = 1e-8  # torch_engine.py:289


def t_spectrum_device(volume: torch.Tensor, pad_factor: float) -> Tuple[torch.Tensor, Tuple[int, int, int]]:
    """complex128 (D, H, W) spectrum of convention T on the volume's device (torch_engine.py:55-88)."""
    d, h, w = volume.shape
    pd, ph, pw = (int(round(n * pad_factor)) for n in (d, h, w))
    vol = torch.zeros((pd, ph, pw), dtype=torch.float64, device=volume.device)
    sd, sh, sw = (pd - d) // 2, (ph - h) // 2, (pw - w) // 2
    vol[sd:sd + d, sh:sh + h, sw:sw + w] = volume.to(torch.float64)
    return torch.fft.fftshift(torch.fft.fftn(vol)), (pd, ph, pw)


def combine_t(mom: np.ndarray, jac: np.ndarray, e_t: np.ndarray, span: float):
    """Per-pose NCC and d NCC / d pose (B, 6) from the fused kernel's moments (float64).

    NCC = A / (sqrt(E E_t) + 1e-8) (torch_engine.py:280-289); the rotation
    derivative uses dA = sum_a Mr_g[a] du_a + Mc_g[a] dv_a and dE = 2 (same with
    the H moments); translations: dA/dtx = 2 pi / span Mr_im, dA/dty = 2 pi / span Mc_im.
    """
    A, E = mom[:, 0], mom[:, 1]
    E = np.maximum(E, 0.0)
    D = np.sqrt(E * e_t) + NCC_EPS
    ncc = A / D
    G = np.einsum("na,naj->nj", mom[:, 4:7], jac[:, 0]) + np.einsum("na,naj->nj", mom[:, 7:10], jac[:, 1])
    Hm = np.einsum("na,naj->nj", mom[:, 10:13], jac[:, 0]) + np.einsum("na,naj->nj", mom[:, 13:16], jac[:, 1])
    safe = E > 0
    dD = np.where(safe[:, None], np.sqrt(e_t)[:, None] * Hm / np.sqrt(np.where(safe, E, 1.0))[:, None], 0.0)
    d = np.zeros((len(A), 6))
    d[:, :3] = G / D[:, None] - (A / D ** 2)[:, None] * dD
    d[:, 3] = 2.0 * math.pi / span * mom[:, 3] / D
    d[:, 4] = 2.0 * math.pi / span * mom[:, 2] / D
    return ncc, d


def check_t_targets(targets, B: int, device) -> int:
    """Validate the targets of a fused convention-T launch before any address is formed from
    them: a complex64 tensor on ``device`` of shape (Bt, R, R) with Bt in {1, B} and rows of
    each target contiguous (stride(-1) == 1, stride(-2) == R); returns R.  The kernel reads
    R * R * 8 bytes per target from ``data_ptr() + k * stride(0)``, so anything else would
    read outside the tensor."""
    if not isinstance(targets, torch.Tensor):
        raise ValueError("targets must be a torch tensor")
    if targets.dtype != torch.complex64:
        raise ValueError(f"targets must be complex64, got {targets.dtype}")
    if targets.device != torch.device(device):
        raise ValueError(f"targets on {targets.device}, expected {device}")
    if targets.dim() != 3 or targets.shape[-1] != targets.shape[-2] or targets.shape[-1] < 2:
        raise ValueError(f"targets must be (Bt, R, R), got {tuple(targets.shape)}")
    if int(targets.shape[0]) not in (1, int(B)):
        raise ValueError(f"targets hold {targets.shape[0]} spectra for {B} poses (expected 1 or {B})")
    R = int(targets.shape[-1])
    if targets.stride(-1) != 1 or targets.stride(-2) != R or (targets.shape[0] > 1 and targets.stride(0) < R * R):
        raise ValueError(f"targets must have contiguous (R, R) rows, got strides {targets.stride()}")
    return R


def check_t_poses(poses: torch.Tensor, device) -> int:
    """(B, 6) contiguous float64 poses on ``device``; returns B."""
    if (not isinstance(poses, torch.Tensor) or poses.dtype != torch.float64 or poses.dim() != 2
            or poses.shape[1] != 6 or not poses.is_contiguous() or poses.shape[0] < 1):
        raise ValueError("poses must be a non-empty contiguous (B, 6) float64 tensor")
    if poses.device != torch.device(device):
        raise ValueError(f"poses on {poses.device}, expected {device}")
    return int(poses.shape[0])


class TSliceEvaluator:
    """Fused convention-T Spectral-NCC (+ exact adjoint) over a batch of poses for one view."""

    def __init__(self, packed: PackedSpectrum, pad_shape, spacing0: float):
        self.packed = packed
        self.pad_shape = tuple(int(s) for s in pad_shape)
        self.spacing0 = float(spacing0)
        self.rows = tile_rows()
        self.ws = _Workspace(packed.device)
        self._energy = None  # (target tensor, its version, key, sum |T|^2 per target)

    def moments(self, poses: np.ndarray, targets: torch.Tensor, view: str, hp_sigma: Optional[float],
                adjoint: bool):
        """(B, 16) float64 moments and (B, 2, 3, 3) Jacobians; ``targets`` (Bt, R, R) complex64, Bt in {1, B}."""
        poses = np.asarray(poses, dtype=np.float64).reshape(-1, 6)
        B = len(poses)
        R = check_t_targets(targets, B, self.packed.device)
        flags = _lib.ITEM_ADJOINT if adjoint else 0
        tmpl = _items.t_template(self.packed, view, R, targets.data_ptr(), flags, self.rows, hp_sigma)
        stride = targets.stride(0) * targets.element_size() if targets.shape[0] > 1 else 0
        ptrs = targets.data_ptr() + stride * np.arange(B, dtype=np.uint64)
        items, jac, n_tiles = _items.t_items(tmpl, self.pad_shape, poses, view, R, self.spacing0, self.rows, ptrs)
        n_chk = _items.validate_items(items, len(self.packed.descs), self.packed.descs, self.rows)
        assert n_chk == n_tiles
        red = self.ws.get("red", B * _lib.NMOM, torch.float64)[: B * _lib.NMOM].view(B, _lib.NMOM)
        launch_items(self.packed, items, n_tiles, int(tmpl["flags"][0]), self.ws, red)
        return red.cpu().numpy(), jac

    def loss_and_grad(self, poses: np.ndarray, targets: torch.Tensor, view: str, hp_sigma: Optional[float],
                      adjoint: bool = True):
        """loss = 1 - mean NCC (torch_engine.py:269-290) and d loss / d poses (B, 6), float64."""
        poses = np.asarray(poses, dtype=np.float64).reshape(-1, 6)
        B = len(poses)
        t = targets.to(torch.complex128)
        e_t = (t.real.square() + t.imag.square()).sum(dim=(-2, -1)).cpu().numpy().reshape(-1)
        e_t = np.broadcast_to(e_t, (B,)) if e_t.size == 1 else e_t
        mom, jac = self.moments(poses, targets, view, hp_sigma, adjoint)
        span = _items.t_phase_span(int(targets.shape[-1]), self.spacing0)
        ncc, d = combine_t(mom, jac, e_t, span)
        return 1.0 - float(np.mean(ncc)), -d / B, ncc


    # -- device path: set-up, fused kernel and combine on the device, no host sync ------
    def _template(self, view: str, R: int, flags: int, hp_sigma: Optional[float]) -> np.ndarray:
        """Validated static item fields for (view, R, flags, hp_sigma) (cached)."""
        key = (view, R, flags, hp_sigma)
        cache = self.__dict__.setdefault("_tmpl", {})
        t = cache.get(key)
        if t is None:
            t = _items.t_template(self.packed, view, R, 1, flags, self.rows, hp_sigma)
            probe, _, _ = _items.t_items(t, self.pad_shape, np.zeros((1, 6)), view, R, self.spacing0, self.rows)
            _items.validate_items(probe, len(self.packed.descs), self.packed.descs, self.rows)
            t = _lib.as_struct(t, _lib.ITEM_DTYPE)
            cache[key] = t
        return t

    def device_loss_and_grad(self, poses: torch.Tensor, targets: torch.Tensor, view: str,
                             hp_sigma: Optional[float], adjoint: bool = True):
        """(ncc (B,), d(1 - mean NCC)/d poses (B, 6)) as float64 device tensors.

        poses: (B, 6) float64 device tensor; targets: (Bt, R, R) complex64 device tensor,
        Bt in {1, B}.  Three stream-ordered launches: fvr_pt_build_items, the fused
        kernel, fvr_pt_combine."""
        B = check_t_poses(poses, self.packed.device)
        R = check_t_targets(targets, B, self.packed.device)
        flags = _lib.ITEM_ADJOINT if adjoint else 0
        tmpl = self._template(view, R, flags, hp_sigma)
        perm = [0, 2, 1] if _items.T_VIEW_LAYOUT[view] == "zxy" else [0, 1, 2]
        order = np.array([[2, 1, 0][a] for a in perm], dtype=np.int32)
        stride = targets.stride(0) * targets.element_size() if targets.shape[0] > 1 else 0
        ntile = int(tmpl["tiles_x"][0]) * _items.tiles_for(R, R, self.rows)[1]
        ws = self.ws
        isz = _lib.ITEM_DTYPE.itemsize
        items = ws.get("t_items", B * isz, torch.uint8)
        jac = ws.get("t_jac", B * 18, torch.float64)
        partials = ws.get("t_partials", B * ntile * _lib.NMOM, torch.float32)
        ncc = ws.get("t_ncc", B, torch.float64)
        grad = torch.empty(6 * B, dtype=torch.float64, device=self.packed.device)  # owned by the caller
        e_t = self.target_energy(targets)
        L = _lib.lib()
        st = _lib.stream_ptr(self.packed.device)
        D, H, W = self.pad_shape
        _lib.check(L.fvr_pt_build_items(_lib.ptr(poses), B, tmpl.ctypes.data, float(D), float(H), float(W),
                                        self.spacing0, R, int(view == "lateral"), order.ctypes.data,
                                        targets.data_ptr(), stride, _lib.ptr(items), _lib.ptr(jac), st),
                   "fvr_pt_build_items")
        variant = self.packed.code | (int(tmpl["flags"][0]) << 4) | (self.rows << 8)
        _lib.check(L.fvr_sncc_eval_v(_lib.ptr(self.packed.table), len(self.packed.descs), _lib.ptr(items), B,
                                     B * ntile, _lib.ptr(partials), variant, st), "fvr_sncc_eval")
        _lib.check(L.fvr_pt_combine(_lib.ptr(items), B, _lib.ptr(partials), _lib.ptr(jac), _lib.ptr(e_t),
                                    1 if e_t.numel() > 1 else 0, _items.t_phase_span(R, self.spacing0), _lib.ptr(ncc),
                                    _lib.ptr(grad), st), "fvr_pt_combine")
        return ncc[:B], grad[: 6 * B].view(B, 6)

    def target_energy(self, targets: torch.Tensor) -> torch.Tensor:
        """sum |T|^2 (float64) of each (R, R) target.  Kept for the target tensor last seen
        (same memory, shape and strides, and no in-place change since: torch's version
        counter, shared with every view of the tensor); the cache holds the tensor, so its
        memory cannot be reused by another one while cached."""
        key = (targets.data_ptr(), tuple(targets.shape), tuple(targets.stride()), targets.dtype, targets.device)
        c = self._energy
        if c is not None and c[2] == key and c[1] == targets._version:
            return c[3]
        t = targets.to(torch.complex128)
        e_t = (t.real.square() + t.imag.square()).sum(dim=(-2, -1)).reshape(-1).contiguous()
        self._energy = (targets, targets._version, key, e_t)
        return e_t

    def device_items(self, B: int) -> np.ndarray:
        raw = self.ws.bufs["t_items"][: B * _lib.ITEM_DTYPE.itemsize].cpu().numpy()
        return raw.view(_lib.ITEM_DTYPE).copy()


class _FusedSpectralNCC(torch.autograd.Function):
    @staticmethod
    def forward(ctx, poses, evaluator, targets, view, hp_sigma):
        dev = evaluator.packed.device
        P = poses.detach().to(dev, torch.float64).contiguous()
        ncc, grad = evaluator.device_loss_and_grad(P, targets, view, hp_sigma, adjoint=bool(ctx.needs_input_grad[0]))
        ctx.grad = grad  # a fresh tensor per call
        loss = torch.empty((), dtype=torch.float32, device=dev)  # 1 - mean NCC (torch_engine.py:290)
        _lib.check(_lib.lib().fvr_mean_loss(_lib.ptr(ncc), int(ncc.numel()), _lib.ptr(loss), _lib.stream_ptr(dev)),
                   "fvr_mean_loss")
        return loss.to(poses.device)

    @staticmethod
    def backward(ctx, g):
        gp = ctx.grad.to(device=g.device, dtype=g.dtype) * g
        return gp, None, None, None, None


class TorchDiffFVREngine(nn.Module):
    """Batched convention-T renderer / Spectral-NCC engine (torch_engine.py:18-337)."""

    def __init__(self, volume: Union[torch.Tensor, np.ndarray], spacing: Tuple[float, float, float] = (1.0, 1.0, 1.0),
                 pad_factor: float = 2.0, device: Optional[Union[str, torch.device]] = None,
                 dtype: torch.dtype = torch.float32, use_triton: bool = True):
        super().__init__()
        if dtype not in _STORAGE_OF:
            raise TypeError(f"dtype must be float32, float16 or bfloat16, got {dtype}")
        vol = torch.as_tensor(volume).float() if not isinstance(volume, torch.Tensor) else volume.float()
        if vol.dim() != 3:
            raise ValueError(f"volume must be 3D (D, H, W), got shape {tuple(vol.shape)}")
        dev = torch.device(device) if device is not None else (vol.device if vol.is_cuda else _lib.default_device())
        if dev.type != "cuda":
            raise RuntimeError("TorchDiffFVREngine runs on a ROCm device only (there is no CPU path)")
        self.device = dev
        self.use_triton = bool(use_triton)
        self.orig_shape = tuple(vol.shape)
        self.spacing = spacing
        self.pad_factor = pad_factor
        spec, self.pad_shape = t_spectrum_device(vol.to(dev), pad_factor)
        self._spec = spec.to(torch.complex64)
        del spec
        self.register_buffer("f_vol_5d", torch.stack([self._spec.real, self._spec.imag], 0).unsqueeze(0).to(dtype))
        self._freq_cache: Dict = {}
        self._hp_cache: Dict = {}
        self._packed: Optional[PackedSpectrum] = None
        self._evaluator: Optional[TSliceEvaluator] = None

    # -- precision / compile -------------------------------------------------
    def to_precision(self, dtype: torch.dtype) -> "TorchDiffFVREngine":
        """Store the spectrum in ``dtype`` (float32, float16, bfloat16; torch_engine.py:90-95).

        The fused path then gathers from FP16/BF16 storage with one power-of-two
        scale per spectrum (no overflow for any spectrum magnitude).
        """
        if dtype not in _STORAGE_OF:
            raise TypeError(f"dtype must be float32, float16 or bfloat16, got {dtype}")
        self.f_vol_5d = torch.stack([self._spec.real, self._spec.imag], 0).unsqueeze(0).to(dtype)
        self._packed = None
        self._evaluator = None
        return self

    def compile_renderer(self, mode: str = "reduce-overhead", fullgraph: bool = False) -> "TorchDiffFVREngine":
        """Accepted for call compatibility (torch_engine.py:97-107); the hot path is already one fused launch."""
        return self

    def _fused(self) -> TSliceEvaluator:
        if self._evaluator is None:
            self._packed = PackedSpectrum(self._spec, _STORAGE_OF[self.f_vol_5d.dtype], layouts=("zyx", "zxy"))
            self._evaluator = TSliceEvaluator(self._packed, self.pad_shape, self.spacing[0])
        return self._evaluator

    # -- grids and filters (torch_engine.py:109-136) ---------------------------
    def _get_freq_grid(self, out_res: int, device, dtype):
        key = (int(out_res), str(device), dtype)
        if key not in self._freq_cache:
            c = torch.linspace(-1.0, 1.0, out_res, device=device, dtype=dtype)
            ku, kv = torch.meshgrid(c, c, indexing="ij")
            self._freq_cache[key] = (ku, kv, ku.reshape(-1).contiguous(), kv.reshape(-1).contiguous())
        return self._freq_cache[key]

    def _get_hp_filter(self, H: int, W: int, cutoff_sigma: float, device, dtype) -> torch.Tensor:
        key = (int(H), int(W), float(cutoff_sigma), str(device), dtype)
        if key not in self._hp_cache:
            y = torch.linspace(-H / 2, H / 2, H, device=device, dtype=dtype)
            x = torch.linspace(-W / 2, W / 2, W, device=device, dtype=dtype)
            yy, xx = torch.meshgrid(y, x, indexing="ij")
            self._hp_cache[key] = (1.0 - torch.exp(-(xx * xx + yy * yy) / (2.0 * cutoff_sigma ** 2)))[None, None]
        return self._hp_cache[key]

    @staticmethod
    def _euler_to_rotation_matrix(yaw_deg, pitch_deg, roll_deg):
        """R = Rz(yaw) Ry(pitch) Rx(roll) (B, 3, 3) (torch_engine.py:138-174)."""
        k = math.pi / 180.0
        a, b, c = yaw_deg * k, pitch_deg * k, roll_deg * k
        o, z = torch.ones_like(a), torch.zeros_like(a)
        ca, sa, cb, sb, cc, sc = torch.cos(a), torch.sin(a), torch.cos(b), torch.sin(b), torch.cos(c), torch.sin(c)

        def m(rows):
            return torch.stack([torch.stack(r, dim=-1) for r in rows], dim=-2)

        Rz = m([[ca, -sa, z], [sa, ca, z], [z, z, o]])
        Ry = m([[cb, z, sb], [z, o, z], [-sb, z, cb]])
        Rx = m([[o, z, z], [z, cc, -sc], [z, sc, cc]])
        return Rz @ (Ry @ Rx)

    # -- slice / DRR (torch_engine.py:176-253) ---------------------------------
    def _extract_slice_spectrum(self, poses: torch.Tensor, view: str = "ap", out_res: int = 128) -> torch.Tensor:
        """Phase-modulated slices (B, out_res, out_res) complex64, differentiable w.r.t. poses."""
        dev = self.device
        p = poses.to(dev)
        B = p.shape[0]
        R = self._euler_to_rotation_matrix(p[:, 0], p[:, 1], p[:, 2])
        if view.lower() == "lateral":
            n90 = torch.full_like(p[:, 0], 90.0)
            R = R @ self._euler_to_rotation_matrix(n90, torch.zeros_like(n90), torch.zeros_like(n90))
        ku, kv, ku_flat, kv_flat = self._get_freq_grid(out_res, dev, torch.float32)
        D, H, W = self.pad_shape
        cxyz = (0.5 * float(max(W - 1, 1)), 0.5 * float(max(H - 1, 1)), 0.5 * float(max(D - 1, 1)))
        scale = torch.tensor(cxyz, device=dev, dtype=R.dtype).view(1, 3)
        s_ri = fourier_slice_3d(self.f_vol_5d, R[:, 0, :] * scale, R[:, 1, :] * scale, ku_flat, kv_flat, cxyz)
        s_ri = s_ri.float().squeeze(0).view(B, 2, out_res, out_res)
        f_slice = torch.complex(s_ri[:, 0], s_ri[:, 1])
        span = out_res * self.spacing[0]
        ang = -2.0 * math.pi * (ku[None] * p[:, 3].view(B, 1, 1) + kv[None] * p[:, 4].view(B, 1, 1)) / (span + 1e-6)
        return f_slice * torch.complex(torch.cos(ang), torch.sin(ang))

    def _forward_core(self, poses: torch.Tensor, view: str = "ap", out_res: int = 128) -> torch.Tensor:
        s = torch.fft.ifftshift(self._extract_slice_spectrum(poses, view, out_res), dim=(-2, -1))
        return torch.relu(torch.real(torch.fft.ifft2(s, dim=(-2, -1)))).unsqueeze(1)

    def forward(self, poses: torch.Tensor, view: str = "ap", out_res: int = 128) -> torch.Tensor:
        """(B, 1, out_res, out_res) DRRs, differentiable w.r.t. poses (torch_engine.py:243-253)."""
        return self._forward_core(poses, view=view, out_res=out_res)

    # -- Spectral-NCC (torch_engine.py:255-337) --------------------------------
    def prepare_spectral_target(self, target_drr: torch.Tensor, high_pass: bool = True,
                                cutoff_sigma: float = 4.0) -> torch.Tensor:
        """fftshift(fft2(target)) [* hp], computed once before optimisation (:255-267)."""
        t = target_drr.to(self.device).float()
        H, W = t.shape[-2:]
        f = torch.fft.fftshift(torch.fft.fft2(t), dim=(-2, -1))
        if high_pass:
            f = f * self._get_hp_filter(H, W, cutoff_sigma, self.device, torch.float32)
        return f

    def _spectral_ncc_core(self, poses: torch.Tensor, f_targ_precomputed: torch.Tensor, view: str = "ap",
                           high_pass: bool = True, cutoff_sigma: float = 4.0) -> torch.Tensor:
        if poses.dim() != 2 or poses.shape[1] != 6:
            raise ValueError(f"poses must be (B, 6), got {tuple(poses.shape)}")
        ft = f_targ_precomputed.to(self.device)
        R = int(ft.shape[-1])
        if ft.shape[-2] != R:
            raise ValueError("f_targ_precomputed must be square (out_res, out_res)")
        ft = ft.reshape(-1, R, R).to(torch.complex64).contiguous()
        if ft.shape[0] not in (1, poses.shape[0]):
            raise ValueError(f"{ft.shape[0]} target spectra for {poses.shape[0]} poses (need 1 or one per pose)")
        sigma = float(cutoff_sigma) if high_pass else None
        return _FusedSpectralNCC.apply(poses, self._fused(), ft, view.lower(), sigma)

    def forward_spectral_ncc(self, poses: torch.Tensor, f_targ_precomputed: torch.Tensor, view: str = "ap",
                             high_pass: bool = True, cutoff_sigma: float = 4.0) -> torch.Tensor:
        """1 - mean_b NCC_b in Fourier space, one fused launch; differentiable w.r.t. poses (:292-307)."""
        return self._spectral_ncc_core(poses, f_targ_precomputed, view, high_pass, cutoff_sigma)

    def spectral_ncc_loss(self, pred_drr: torch.Tensor, target_drr: torch.Tensor, high_pass: bool = True,
                          cutoff_sigma: float = 4.0) -> torch.Tensor:
        """Image-space Spectral-NCC loss 1 - mean NCC(fft2(pred), fft2(target)) (:309-337)."""
        H, W = pred_drr.shape[-2:]
        Fp = torch.fft.fftshift(torch.fft.fft2(pred_drr), dim=(-2, -1))
        Ft = torch.fft.fftshift(torch.fft.fft2(target_drr.to(pred_drr.device)), dim=(-2, -1))
        if high_pass:
            hp = self._get_hp_filter(H, W, cutoff_sigma, pred_drr.device, pred_drr.dtype)
            Fp, Ft = Fp * hp, Ft * hp
        cross = torch.sum(torch.real(Fp * torch.conj(Ft)), dim=(-2, -1))
        ep = torch.sum(Fp.real.square() + Fp.imag.square(), dim=(-2, -1))
        et = torch.sum(Ft.real.square() + Ft.imag.square(), dim=(-2, -1))
        return 1.0 - torch.mean(cross / (torch.sqrt(ep * et) + NCC_EPS))
