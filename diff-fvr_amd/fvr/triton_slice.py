"""The reference's import path for its slice operator (``fvr.triton_slice``).

``fourier_slice_3d`` and ``triton_slice_available`` are the HIP-backed
versions from ``slice_ops``; the name is kept so that code importing
``fvr.triton_slice.fourier_slice_3d`` (triton_slice.py:648-661) runs unchanged.
"""

from .slice_ops import fourier_slice_3d, triton_slice_available

__all__ = ["fourier_slice_3d", "triton_slice_available"]
