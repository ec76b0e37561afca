"""MI355X-native Diff-FVR rendering hot path.

Drop-in replacements for the reference package's (``pareelamre/diff-fvr``,
``src/fvr/__init__.py:30-50``) rendering and registration names that sit on
the hot path; every numerical step runs in the HIP kernels of
``libfvr_hip.so``.  Importing the package does not touch the GPU; using a
renderer without a ROCm device raises.
"""

from . import _lib
from .diff_fvr import DifferentiableFourierVolumeRenderer
from .gradients import SpectralGradientEngine
from .gradient_optimizer import GradientRegistrationOptimizer
from .core_engine import RobustDiffFVREngine, detect_radio_opaque_metal
from .registration_3d2d import compute_mtre_6dof
from .phantom import ellipsoid_phantom
from .triton_slice import fourier_slice_3d
from .torch_engine import TorchDiffFVREngine
from .cone_beam_fvr import (PerspectiveSlabFVR, SLAB_PRESETS, edge_dominant_ncc, image_ncc, pou_pinhole_warp_sum,
                            sample_isocenter_parity_offsets)

__version__ = "0.1.0"

__all__ = [
    "DifferentiableFourierVolumeRenderer",
    "SpectralGradientEngine",
    "GradientRegistrationOptimizer",
    "RobustDiffFVREngine",
    "detect_radio_opaque_metal",
    "compute_mtre_6dof",
    "ellipsoid_phantom",
    "fourier_slice_3d",
    "TorchDiffFVREngine",
    "PerspectiveSlabFVR",
    "SLAB_PRESETS",
    "sample_isocenter_parity_offsets",
    "edge_dominant_ncc",
    "image_ncc",
    "pou_pinhole_warp_sum",
]
