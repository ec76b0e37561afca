"""Parallel-beam Differentiable Fourier Volume Rendering (convention N) on MI355X.

Drop-in for ``src/fvr/diff_fvr.py:7-195`` (``DifferentiableFourierVolumeRenderer``):
same constructor, attributes (``pz, py, px, orig_shape, pad_factor,
apply_premultiplication, f_3d``) and methods, with NumPy in and NumPy out.
The spectrum lives on the GPU (float64 precompute, then the packed
complex64 / FP16 / BF16 layouts), and every slice is sampled by the HIP
kernel in ``csrc/fvr_sncc.hip`` with scipy's order-1 'constant' interpolation
rule (``diff_fvr.py:141-142``).  There is no CPU path.
"""

from __future__ import annotations

from typing import Dict, Optional, Tuple

import numpy as np
import torch

from . import _items, _lib
from .sncc import _Workspace, pair_moments, run_items, tile_rows
from .spectrum import PackedSpectrum, padded_shape, padded_spectrum

# zero-margin copies (single slices, FD scheme) and wrapped copies (Hermitian-fold batches)
PACK_LAYOUTS = ("zyx", "zxy", "zyx_h", "zxy_h")


class DifferentiableFourierVolumeRenderer:
    """Differentiable & spectral Fourier volume renderer (parallel beam).

    Extra keyword arguments over the reference: ``storage`` ("f32", "fp16",
    "bf16": spectrum storage of the HIP layouts; interpolation and sums are
    FP32 in all cases) and ``device``.
    """

    def __init__(self, volume: np.ndarray, pad_factor: float = 2.0, apply_premultiplication: bool = True,
                 storage: str = "f32", device: Optional[torch.device] = None):
        self.device = torch.device(device) if device is not None else _lib.default_device()
        vol = np.asarray(volume)
        self.orig_shape = vol.shape
        self.pad_factor = pad_factor
        self.apply_premultiplication = apply_premultiplication
        self.pz, self.py, self.px = padded_shape(vol.shape, pad_factor)
        spec = padded_spectrum(vol, pad_factor, apply_premultiplication, self.device)
        self._spec = spec  # complex128, centred, (pz, py, px)
        self.storage = storage
        self.packed = PackedSpectrum(spec, dtype=storage, layouts=PACK_LAYOUTS)
        self._ws = _Workspace(self.device)
        self._zero_targets: Dict[Tuple[int, int], torch.Tensor] = {}

    # -- attributes ---------------------------------------------------------
    @property
    def f_3d(self) -> np.ndarray:
        """The centred complex128 spectrum (copied to the host on access)."""
        return self._spec.cpu().numpy()

    def repack(self, storage: str) -> "DifferentiableFourierVolumeRenderer":
        """Re-store the spectrum as "f32", "fp16" or "bf16" (one power-of-two scale)."""
        self.storage = storage
        self.packed = PackedSpectrum(self._spec, dtype=storage, layouts=PACK_LAYOUTS)
        return self

    # -- geometry (diff_fvr.py:56-104) -------------------------------------
    def _euler_to_basis(self, yaw_deg: float, pitch_deg: float, roll_deg: float, view: str = "lateral"):
        return _items.euler_to_basis_n(yaw_deg, pitch_deg, roll_deg, view)

    def _slice_dims(self, view: str) -> Tuple[int, int]:
        return _items.slice_dims_n((self.pz, self.py, self.px), view)

    def _zero_target(self, h: int, w: int) -> torch.Tensor:
        t = self._zero_targets.get((h, w))
        if t is None:
            t = torch.zeros((h, w), dtype=torch.complex64, device=self.device)
            self._zero_targets[(h, w)] = t
        return t

    # -- slices --------------------------------------------------------------
    def extract_frequency_slice_device(self, yaw_deg=0.0, pitch_deg=0.0, roll_deg=0.0, tx=0.0, ty=0.0,
                                       view="lateral", interp_order=1) -> torch.Tensor:
        """Phase-modulated slice as a complex64 (slice_h, slice_w) device tensor."""
        if interp_order != 1:
            raise NotImplementedError("the HIP slice kernel implements trilinear (interp_order=1) sampling only")
        h, w = self._slice_dims(view)
        out = torch.empty((h, w), dtype=torch.complex64, device=self.device)
        it = _items.n_item(self.packed, view, float(yaw_deg), float(pitch_deg), float(roll_deg), float(tx),
                           float(ty), self._zero_target(h, w).data_ptr(), _lib.ITEM_WRITE_SLICE, 0, tile_rows(),
                           slice_out=out.data_ptr())
        run_items(self.packed, it, _lib.ITEM_WRITE_SLICE, self._ws)
        return out

    def extract_frequency_slice(self, yaw_deg: float = 0.0, pitch_deg: float = 0.0, roll_deg: float = 0.0,
                                tx: float = 0.0, ty: float = 0.0, view: str = "lateral",
                                interp_order: int = 1) -> np.ndarray:
        """Parallel-beam frequency slice (tx/ty in in-plane voxels) as complex128 NumPy."""
        s = self.extract_frequency_slice_device(yaw_deg, pitch_deg, roll_deg, tx, ty, view, interp_order)
        return s.cpu().numpy().astype(np.complex128)

    def render_drr_device(self, yaw_deg=0.0, pitch_deg=0.0, roll_deg=0.0, tx=0.0, ty=0.0, view="lateral",
                          interp_order=1, crop_to_roi=True) -> torch.Tensor:
        """DRR (diff_fvr.py:152-182) as a float64 device tensor."""
        s = self.extract_frequency_slice_device(yaw_deg, pitch_deg, roll_deg, tx, ty, view, interp_order)
        slice_h, slice_w = s.shape
        proj = torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(s.to(torch.complex128))))
        scale = (self.pz * self.py * self.px) / (slice_h * slice_w)
        drr = torch.clamp(proj.real * scale, min=0.0)
        if crop_to_roi:
            nz, ny, nx = self.orig_shape
            th = nz if view in ("lateral", "ap") else ny
            tw = ny if view == "lateral" else nx
            sh, sw = (slice_h - th) // 2, (slice_w - tw) // 2
            drr = drr[sh:sh + th, sw:sw + tw]
        return drr

    def render_drr(self, yaw_deg: float = 0.0, pitch_deg: float = 0.0, roll_deg: float = 0.0, tx: float = 0.0,
                   ty: float = 0.0, view: str = "lateral", interp_order: int = 1,
                   crop_to_roi: bool = True) -> np.ndarray:
        return self.render_drr_device(yaw_deg, pitch_deg, roll_deg, tx, ty, view, interp_order,
                                      crop_to_roi).cpu().numpy()

    # -- loss (diff_fvr.py:184-195) ------------------------------------------
    def compute_spectral_ncc(self, f_slice_candidate, f_target_precomputed) -> float:
        """Re sum(S conj T) / (||S|| ||T||) on the device; 0 if a norm is 0."""
        a = _as_device_c64(f_slice_candidate, self.device)
        b = _as_device_c64(f_target_precomputed, self.device)
        m = pair_moments(a, b)
        ns, nt = np.sqrt(m[1]), np.sqrt(m[2])
        if ns * nt == 0:
            return 0.0
        return float(m[0] / (ns * nt))


def _as_device_c64(x, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        _lib.require_device(x, "spectral input")
        return x.to(torch.complex64)
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex64))).to(device)
