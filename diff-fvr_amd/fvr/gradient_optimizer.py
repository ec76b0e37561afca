"""6-DoF biplanar registration by Adam: the reference's ``register_adam`` interface
(``src/fvr/gradient_optimizer.py:15-90``) over this build's device pieces.

Each iteration evaluates the biplanar loss and gradient with the HIP engine (the
reference's 14-slice finite-difference scheme by default, or the exact adjoint)
and takes the Adam step with the device kernel ``fvr_adam_step`` (``DeviceAdam``
with one pose): the reference's hyper-parameters (beta 0.9 / 0.999, eps 1e-8,
separate rotation / translation rates, poses clipped to +-20 deg / +-25, best-loss
tracking, stop once the largest step is below ``tol``) are that kernel's defaults.
The host reads one flag per iteration to stop, and the pose for the trajectory.
"""

from __future__ import annotations

import time
from typing import Any, Dict

import numpy as np
import torch

from .gradients import SpectralGradientEngine
from .multistart import DeviceAdam


class GradientRegistrationOptimizer:
    """Rapid 6-DoF biplanar registration using Fourier-domain gradients."""

    def __init__(self, renderer):
        self.renderer = renderer
        self.grad_engine = SpectralGradientEngine(renderer)

    def register_adam(self, f_target_ap: np.ndarray, f_target_lat: np.ndarray, init_pose: np.ndarray,
                      lr_rot: float = 0.25, lr_trans: float = 0.35, max_iter: int = 40, tol: float = 1e-4,
                      rotation_grad: str = "fd") -> Dict[str, Any]:
        """Adam from ``init_pose`` (yaw, pitch, roll in degrees; tx, ty, tz in voxels).

        Returns the reference's dict: success (best loss < 0.05), final_pose (the pose of
        the lowest loss seen), final_loss, iterations, evaluations (7 per iteration, the
        reference's count), time_sec and trajectory [(iteration, loss, pose before the
        step)]."""
        started = time.time()
        dev = self.renderer.device
        f64 = dict(dtype=torch.float64, device=dev)
        pose = torch.as_tensor(np.asarray(init_pose, dtype=np.float64).reshape(1, 6)).to(dev).contiguous()
        adam = DeviceAdam(1, dev, lr_rot=lr_rot, lr_trans=lr_trans, tol=tol)
        adam.reset(pose)
        loss_dev = torch.zeros(1, **f64)
        grad_dev = torch.zeros((1, 6), **f64)
        trajectory = []
        iterations = 0
        for it in range(1, int(max_iter) + 1):
            iterations = it
            here = pose[0].cpu().numpy().copy()
            loss, grad = self.grad_engine.compute_biplanar_loss_and_gradients(here, f_target_ap, f_target_lat,
                                                                              rotation_grad=rotation_grad)
            trajectory.append((it, float(loss), here))
            loss_dev.copy_(torch.tensor([float(loss)], dtype=torch.float64))
            grad_dev.copy_(torch.as_tensor(np.asarray(grad, dtype=np.float64).reshape(1, 6)))
            adam.step(pose, loss_dev, grad_dev)
            if int(adam.done[0]):
                break
        best_loss = float(adam.best_loss[0])
        return {
            "success": bool(best_loss < 0.05),
            "final_pose": adam.best_pose[0].cpu().numpy(),
            "final_loss": best_loss,
            "iterations": iterations,
            "evaluations": iterations * 7,
            "time_sec": time.time() - started,
            "trajectory": trajectory,
        }
