"""The reference's import path for its slice operator (``fvr.triton_slice``).

``fourier_slice_3d``, ``pou_pinhole_warp_sum`` and ``triton_slice_available``
are the HIP-backed versions; the module name is kept so that code importing
them from ``fvr.triton_slice`` (triton_slice.py:648-661, 787-808) runs unchanged.
"""

from .cone_beam_fvr import pou_pinhole_warp_sum
from .slice_ops import fourier_slice_3d, triton_slice_available

__all__ = ["fourier_slice_3d", "pou_pinhole_warp_sum", "triton_slice_available"]
