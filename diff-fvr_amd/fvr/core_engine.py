"""mm-unit parallel-beam registration engine (drop-in for ``src/fvr/core_engine.py:30-224``).

Input validation, target-spectrum preparation (``_prepare_spectrum``,
core_engine.py:71-85) and the Spectral-NCC biplanar objective
(core_engine.py:195-205) run on the GPU: each Powell evaluation of the clean
objective is one fused loss-only launch over both views (``BiplanarEvaluator``
without the adjoint).  The Powell driver (``scipy.optimize.minimize``, one pose
per call, as in the reference) and the metal-masked spatial loss are host logic
around GPU-rendered DRRs.
"""

from __future__ import annotations

import time
from typing import Any, Dict, Optional, Tuple

import numpy as np
import torch
from scipy.optimize import minimize

from .diff_fvr import DifferentiableFourierVolumeRenderer
from .powell import batched_powell
from .sncc import BiplanarEvaluator


def detect_radio_opaque_metal(img: np.ndarray, saturation_threshold: float = 0.98) -> np.ndarray:
    """Saturated plateaus of the normalised image (core_engine.py:21-27)."""
    norm = (img - img.min()) / (img.max() - img.min() + 1e-8)
    return norm >= saturation_threshold


def prepare_spectrum_device(img, slice_h: int, slice_w: int, device) -> torch.Tensor:
    """Centred zero-pad into (slice_h, slice_w), fftshift(fft2(ifftshift(.))) (core_engine.py:71-85)."""
    im = torch.as_tensor(img if isinstance(img, torch.Tensor) else np.asarray(img, dtype=np.float64), device=device)
    im = im.to(torch.float64)
    th, tw = im.shape
    padded = torch.zeros((slice_h, slice_w), dtype=torch.float64, device=device)
    sh, sw = (slice_h - th) // 2, (slice_w - tw) // 2
    padded[sh:sh + th, sw:sw + tw] = im
    return torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(padded)))


class RobustDiffFVREngine:
    """Parallel-beam registration on a cubic, isotropic volume; poses (deg, mm)."""

    def __init__(self, volume: np.ndarray, spacing: Tuple[float, float, float] = (1.0, 1.0, 1.0),
                 pad_factor: float = 2.0, sdd_mm: Optional[float] = None, sad_mm: Optional[float] = None,
                 storage: str = "f32", device=None):
        if sdd_mm is not None or sad_mm is not None:
            raise ValueError("Only parallel-beam geometry is supported; SDD/SAD calibration is not implemented.")
        spacing = np.asarray(spacing, dtype=float)
        if spacing.shape != (3,) or not np.all(np.isfinite(spacing)) or np.any(spacing <= 0):
            raise ValueError("spacing must contain three finite positive values in mm (Z, Y, X).")
        if not np.allclose(spacing, spacing[0], rtol=1e-6, atol=0):
            raise ValueError("Resample the volume to isotropic spacing before registration.")
        if volume.ndim != 3 or len(set(volume.shape)) != 1 or min(volume.shape) < 2:
            raise ValueError("Use a cubic 3D volume with at least two voxels per axis.")
        if not np.all(np.isfinite(volume)):
            raise ValueError("volume must contain only finite values.")
        self.volume = volume.astype(np.float64)
        self.spacing = tuple(spacing)
        self.renderer = DifferentiableFourierVolumeRenderer(self.volume, pad_factor=pad_factor,
                                                            apply_premultiplication=True, storage=storage,
                                                            device=device)

    def detect_metal_hardware_mask(self, img: np.ndarray, threshold: float = 0.98) -> np.ndarray:
        return ~detect_radio_opaque_metal(img, saturation_threshold=threshold)

    def _prepare_spectrum_device(self, img, view: str) -> torch.Tensor:
        slice_h, slice_w = self.renderer._slice_dims(view)
        return prepare_spectrum_device(img, slice_h, slice_w, self.renderer.device)

    def _prepare_spectrum(self, img: np.ndarray, view: str) -> np.ndarray:
        return self._prepare_spectrum_device(img, view).cpu().numpy()

    def render_calibrated_drr(self, yaw: float = 0.0, pitch: float = 0.0, roll: float = 0.0, tx: float = 0.0,
                              ty: float = 0.0, view: str = "lateral", crop_to_roi: bool = True) -> np.ndarray:
        """Parallel-beam DRR; tx/ty are in-plane translations in mm (core_engine.py:87-107)."""
        if view not in ("ap", "lateral", "axial"):
            raise ValueError("view must be ap, lateral or axial.")
        return self.renderer.render_drr(yaw_deg=yaw, pitch_deg=pitch, roll_deg=roll, tx=tx / self.spacing[0],
                                        ty=ty / self.spacing[0], view=view, crop_to_roi=crop_to_roi)

    def register_biplanar_robust(self, target_ap: np.ndarray, target_lat: np.ndarray,
                                 initial_pose=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0), handle_metal: bool = True,
                                 max_iter: int = 50, bounds_trans_mm: float = 25.0,
                                 bounds_rot_deg: float = 20.0, optimizer: str = "powell") -> Dict[str, Any]:
        """Powell over Spectral-NCC (clean) or masked spatial NCC (metal), poses in (deg, mm).

        ``optimizer="powell"`` is the reference's scipy Powell (one pose per evaluation);
        ``"batched_powell"`` runs ``fvr.powell.batched_powell`` on the clean objective: every
        line-search round evaluates a grid of poses along the direction in one fused
        loss-only launch (the metal-masked objective always uses scipy's Powell)."""
        if optimizer not in ("powell", "batched_powell"):
            raise ValueError(f"optimizer must be 'powell' or 'batched_powell', got {optimizer!r}")
        for name, target in (("AP", target_ap), ("lateral", target_lat)):
            if target.shape != self.volume.shape[:2] or not np.all(np.isfinite(target)):
                raise ValueError(f"{name} target must be finite with shape {self.volume.shape[:2]}; "
                                 "match the parallel-beam projection grid before registration.")
            if np.ptp(target) == 0:
                raise ValueError(f"{name} target must contain intensity variation for registration.")
        initial_pose = np.asarray(initial_pose, dtype=float)
        if initial_pose.shape != (6,) or not np.all(np.isfinite(initial_pose)):
            raise ValueError("initial_pose must contain six finite values.")
        if not np.all(np.isfinite([bounds_trans_mm, bounds_rot_deg])) or min(bounds_trans_mm, bounds_rot_deg) <= 0:
            raise ValueError("Translation and rotation bounds must be finite and positive.")
        if max_iter < 1:
            raise ValueError("max_iter must be positive.")
        t0 = time.time()
        eval_count = 0
        mask_ap = detect_radio_opaque_metal(target_ap) if handle_metal else np.zeros_like(target_ap, dtype=bool)
        mask_lat = detect_radio_opaque_metal(target_lat) if handle_metal else np.zeros_like(target_lat, dtype=bool)
        metal_present = bool((np.mean(mask_ap) > 0.03) or (np.mean(mask_lat) > 0.03))
        bounds = [(-bounds_rot_deg, bounds_rot_deg)] * 3 + [(-bounds_trans_mm, bounds_trans_mm)] * 3
        r = self.renderer
        sp = self.spacing[0]
        if metal_present:
            valid_ap, valid_lat = ~mask_ap, ~mask_lat
            ta, tl = target_ap[valid_ap], target_lat[valid_lat]
            t_ap = (ta - ta.mean()) / (ta.std() + 1e-8)
            t_lat = (tl - tl.mean()) / (tl.std() + 1e-8)

            def objective(p):
                nonlocal eval_count
                eval_count += 1
                yaw, pitch, roll, tx, ty, tz = p
                tx, ty, tz = np.array([tx, ty, tz]) / sp
                d_ap = r.render_drr(yaw, pitch, roll, tx=tx, ty=tz, view="ap")[valid_ap]
                d_lat = r.render_drr(yaw, pitch, roll, tx=ty, ty=tz, view="lateral")[valid_lat]
                ncc_ap = np.mean((d_ap - d_ap.mean()) / (d_ap.std() + 1e-8) * t_ap)
                ncc_lat = np.mean((d_lat - d_lat.mean()) / (d_lat.std() + 1e-8) * t_lat)
                return float(0.5 * (1.0 - ncc_ap) + 0.5 * (1.0 - ncc_lat))
        else:
            f_ap = self._prepare_spectrum_device(target_ap, "ap").to(torch.complex64)
            f_lat = self._prepare_spectrum_device(target_lat, "lateral").to(torch.complex64)
            # one fused loss-only launch per evaluation: both views' slices, phase ramps and
            # Spectral-NCC sums (the slices themselves are never stored)
            ev = BiplanarEvaluator(r.packed, f_ap, f_lat, adjoint=False)
            pose = np.zeros((1, 6))

            def objective(p):
                nonlocal eval_count
                eval_count += 1
                pose[0, :3] = p[:3]
                pose[0, 3:] = np.asarray(p[3:], dtype=float) / sp  # (tx, ty, tz) mm -> voxels
                loss, _ = ev.evaluate(pose)
                return float(loss[0])

            if optimizer == "batched_powell":
                def loss_batch(P):
                    nonlocal eval_count
                    Q = np.array(P, dtype=np.float64).reshape(-1, 6)
                    eval_count += len(Q)
                    Q[:, 3:] /= sp
                    return ev.evaluate(Q)[0]

                res = batched_powell(loss_batch, np.array(initial_pose), bounds, max_iter=max_iter, ftol=1e-4)
                return {
                    "success": bool(res["success"]),
                    "final_pose": res["x"],
                    "final_loss": float(res["fun"]),
                    "evaluations": eval_count,
                    "launches": res["calls"],
                    "metal_detected": metal_present,
                    "time_sec": time.time() - t0,
                }

        res = minimize(objective, x0=np.array(initial_pose), method="Powell", bounds=bounds,
                       options={"maxiter": max_iter, "ftol": 1e-4})
        return {
            "success": bool(res.success),
            "final_pose": res.x,
            "final_loss": float(res.fun),
            "evaluations": eval_count,
            "launches": eval_count,
            "metal_detected": metal_present,
            "time_sec": time.time() - t0,
        }
