"""Centred 3D spectra on the device and their packed HIP layouts.

``padded_spectrum`` restates the reference's precompute
(``diff_fvr.py:22-54``: sinc^2 premultiplication, centred zero padding,
``fftshift(fftn(ifftshift(.)))``) in float64 on the GPU (rocFFT through
torch.fft).  ``PackedSpectrum`` stores it in the layouts the fused kernel
gathers from (``fvr_spec_pack``): one copy whose fast axis is x (AP and
axial views) and one whose fast axis is y (lateral view), each with zero
margins, in complex64 or FP16/BF16 with one power-of-two scale per spectrum.
"""

from __future__ import annotations

import ctypes
from typing import Dict, Sequence, Tuple

import numpy as np
import torch

from . import _lib


def sinc2_correction(n_orig: int, n_pad: int, device, dtype=torch.float64) -> torch.Tensor:
    """1 / max(sinc^2((i - n/2)/p), 0.05)  (diff_fvr.py:35-39)."""
    c = (torch.arange(n_orig, device=device, dtype=dtype) - n_orig / 2.0) / n_pad
    return 1.0 / torch.clamp(torch.sinc(c) ** 2, min=0.05)


def padded_shape(shape: Sequence[int], pad_factor: float) -> Tuple[int, int, int]:
    return tuple(int(round(int(n) * pad_factor)) for n in shape)


def padded_spectrum(volume, pad_factor: float = 2.0, premultiply: bool = True,
                    device: torch.device | None = None) -> torch.Tensor:
    """Convention-N spectrum (complex128, index (z, y, x)) on the device."""
    device = device or _lib.default_device()
    vol = torch.as_tensor(np.asarray(volume) if not isinstance(volume, torch.Tensor) else volume)
    vol = vol.to(device=device, dtype=torch.float64)
    nz, ny, nx = vol.shape
    pz, py, px = padded_shape(vol.shape, pad_factor)
    if premultiply:
        vol = vol * (sinc2_correction(nz, pz, device)[:, None, None]
                     * sinc2_correction(ny, py, device)[None, :, None]
                     * sinc2_correction(nx, px, device)[None, None, :])
    padded = torch.zeros((pz, py, px), dtype=torch.float64, device=device)
    sz, sy, sx = (pz - nz) // 2, (py - ny) // 2, (px - nx) // 2
    padded[sz:sz + nz, sy:sy + ny, sx:sx + nx] = vol
    del vol
    spec = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(padded)))
    return spec


class PackedSpectrum:
    """Device-resident packed layouts of one centred complex spectrum.

    ``layouts`` names which copies to build: ``"zyx"`` (fast axis x) and/or
    ``"zxy"`` (fast axis y), and their wrapped forms ``"zyx_h"`` / ``"zxy_h"``
    (index n of every axis holds the periodic copy of index 0, for the
    Hermitian-fold kernel).  ``table`` is the device array of ``fvr_layout``
    descriptors in that order; ``index[name]`` its position.
    """

    def __init__(self, spec: torch.Tensor, dtype: str = "f32", layouts: Sequence[str] = ("zyx", "zxy")):
        _lib.require_device(spec, "PackedSpectrum")
        if dtype not in _lib.DTYPE_CODES:
            raise ValueError(f"dtype must be one of {tuple(_lib.DTYPE_CODES)}, got {dtype!r}")
        L = _lib.lib()
        self.device = spec.device
        self.dtype = dtype
        self.code = _lib.DTYPE_CODES[dtype]
        self.shape = tuple(int(s) for s in spec.shape)  # (Z, Y, X)
        src = spec.to(torch.complex64).contiguous()
        self.index: Dict[str, int] = {}
        self.storage: Dict[str, torch.Tensor] = {}
        self.scales: Dict[str, torch.Tensor] = {}
        descs = np.zeros(len(layouts), dtype=_lib.LAYOUT_DTYPE)
        Z, Y, X = self.shape
        for k, name in enumerate(layouts):
            if name not in ("zyx", "zxy", "zyx_h", "zxy_h"):
                raise ValueError(f"unknown layout {name!r}")
            transpose = name.startswith("zxy")
            wrap = name.endswith("_h")
            n = np.array([Z, X, Y] if transpose else [Z, Y, X], dtype=np.int32)
            one = np.zeros(1, dtype=_lib.LAYOUT_DTYPE)
            nbytes, nsc = ctypes.c_uint64(0), ctypes.c_uint64(0)
            code = self.code | ((_lib.LAYOUT_WRAP << 8) if wrap else 0)
            _lib.check(L.fvr_layout_init(one.ctypes.data, n.ctypes.data, code, ctypes.byref(nbytes),
                                         ctypes.byref(nsc)), "fvr_layout_init")
            store = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
            one["base"] = store.data_ptr()
            if nsc.value:
                sc = torch.empty(int(nsc.value), dtype=torch.float32, device=self.device)
                one["scales"] = sc.data_ptr()
                self.scales[name] = sc
            _lib.check(L.fvr_spec_pack(one.ctypes.data, src.data_ptr(), Z, Y, X, int(transpose),
                                       _lib.stream_ptr(self.device)), "fvr_spec_pack")
            self.storage[name] = store
            self.index[name] = k
            descs[k] = one[0]
        torch.cuda.current_stream(self.device).synchronize()  # src is freed on return
        self.descs = descs
        self.table = _lib.struct_to_device(descs, self.device, _lib.LAYOUT_DTYPE)

    def layout_for(self, view: str, herm: bool = False) -> Tuple[int, Tuple[int, int, int]]:
        """(layout index, permutation layout-axis -> (z, y, x) axis) for a view
        (the wrapped copy when ``herm``)."""
        sfx = "_h" if herm else ""
        if view == "lateral" and "zxy" + sfx in self.index:
            return self.index["zxy" + sfx], (0, 2, 1)
        return self.index["zyx" + sfx], (0, 1, 2)

    def has_herm(self) -> bool:
        return "zyx_h" in self.index

    def desc(self, k: int):
        return self.descs[k]

    def nbytes(self) -> int:
        return sum(int(t.numel()) for t in self.storage.values()) + sum(4 * int(t.numel()) for t in self.scales.values())
