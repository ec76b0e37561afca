"""6-DoF biplanar registration by Adam over the HIP loss/gradient.

Drop-in for ``src/fvr/gradient_optimizer.py:15-90``: the same Adam update
(beta 0.9/0.999, eps 1e-8, decoupled rotation/translation learning rates,
clipping to +-20 deg / +-25, early stop when max|step| < tol) on the host,
with every loss/gradient evaluation on the GPU.
"""

from __future__ import annotations

import time
from typing import Any, Dict

import numpy as np

from .gradients import SpectralGradientEngine


class GradientRegistrationOptimizer:
    """Rapid 6-DoF biplanar registration using Fourier-domain gradients."""

    def __init__(self, renderer):
        self.renderer = renderer
        self.grad_engine = SpectralGradientEngine(renderer)

    def register_adam(self, f_target_ap: np.ndarray, f_target_lat: np.ndarray, init_pose: np.ndarray,
                      lr_rot: float = 0.25, lr_trans: float = 0.35, max_iter: int = 40, tol: float = 1e-4,
                      rotation_grad: str = "fd") -> Dict[str, Any]:
        """Adam with decoupled rotation and translation learning rates (gradient_optimizer.py:24-90)."""
        t0 = time.time()
        pose = np.array(init_pose, dtype=np.float64)
        m = np.zeros(6)
        v = np.zeros(6)
        beta1, beta2, eps = 0.9, 0.999, 1e-8
        lr_vec = np.array([lr_rot] * 3 + [lr_trans] * 3)
        trajectory = []
        best_loss, best_pose = float("inf"), pose.copy()
        step = 0
        for step in range(1, max_iter + 1):
            loss, grad = self.grad_engine.compute_biplanar_loss_and_gradients(
                pose, f_target_ap, f_target_lat, rotation_grad=rotation_grad)
            trajectory.append((step, loss, pose.copy()))
            if loss < best_loss:
                best_loss, best_pose = loss, pose.copy()
            m = beta1 * m + (1.0 - beta1) * grad
            v = beta2 * v + (1.0 - beta2) * grad ** 2
            m_hat = m / (1.0 - beta1 ** step)
            v_hat = v / (1.0 - beta2 ** step)
            step_update = lr_vec * (m_hat / (np.sqrt(v_hat) + eps))
            pose = pose - step_update
            pose[:3] = np.clip(pose[:3], -20.0, 20.0)
            pose[3:] = np.clip(pose[3:], -25.0, 25.0)
            if np.max(np.abs(step_update)) < tol:
                break
        return {
            "success": bool(best_loss < 0.05),
            "final_pose": best_pose,
            "final_loss": float(best_loss),
            "iterations": step,
            "evaluations": step * 7,
            "time_sec": time.time() - t0,
            "trajectory": trajectory,
        }
