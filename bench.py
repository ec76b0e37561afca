#!/usr/bin/env python3
"""Benchmark of the Diff-FVR hot path on one MI355X.

Workload: convention-N biplanar Spectral-NCC forward loss + exact 6-DoF pose
gradient, seeded 256^3 ellipsoid phantom (1 mm voxels, 12 ellipsoids,
mu ~ U(0.01, 0.05)/mm) zero-padded 2x to a 512^3 spectrum, AP + lateral
views (512^2 slice samples each), 64 candidate poses (rotations U(-15,15) deg,
translations U(-20,20) mm) per launch, against AP/LAT target spectra of 256^2
DRRs rendered at a pose with rotations and translations U(-5,5).

One step = two stream-ordered launches for the whole batch: one fused HIP
launch over all 2x64 (pose, view) items, then one launch that does the
per-pose fixed-order float64 reduction + biplanar combine (loss and exact
gradient) and the float64 set-up of the items for the next step (spectrum,
targets and poses resident in HBM; loss and gradient left in HBM, no host sync
inside a step); the first step's items are built before the timed region.  The
three-launch form and a HIP-graph replay of the step are reported beside it.
The launch plan of the fused kernel is autotuned first.  `value` is
pose evaluations per second of the batched FP16-spectrum form (FP32
interpolation and sums); the FP32-spectrum batched form, the plain
one-pose-per-launch FP32 form, the PoU-FVR perspective workload, the
convention-T engine and the ragged multistart loop are reported beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diff-fvr_amd"))
sys.path.insert(0, ROOT)

METRIC = "pose evals/sec (fwd+grad) at 256^3 spectrum, 64-cand batch; % HBM roofline"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(size: int, n_cand: int):
    import numpy as np
    from fvr.phantom import ellipsoid_phantom, random_poses
    vol = ellipsoid_phantom(size, n_ellipsoids=12, seed=1234)
    rng = np.random.default_rng(2345)
    true_pose = np.concatenate([rng.uniform(-5, 5, 3), rng.uniform(-5, 5, 3)])
    cands = random_poses(n_cand, 15.0, 20.0, seed=3456)
    return vol, true_pose, cands


def make_targets(renderer, true_pose):
    """AP / LAT target spectra: 256^2 DRRs at the true pose, re-padded (core_engine.py:71-85)."""
    from fvr.core_engine import prepare_spectrum_device
    yaw, pitch, roll, tx, ty, tz = [float(v) for v in true_pose]
    out = []
    for view, a in (("ap", tx), ("lateral", ty)):
        img = renderer.render_drr_device(yaw, pitch, roll, tx=a, ty=tz, view=view, crop_to_roi=True)
        h, w = renderer._slice_dims(view)
        out.append(prepare_spectrum_device(img, h, w, renderer.device).to(__import__("torch").complex64))
    return out


KERNEL_EVENT_EVERY = 8
KERNEL_SAMPLES = 64


def time_batched(ev, P, steps: int, warmup: int):
    """Wall time per step -- the fused kernel over the batch's items, then fvr_pn_step
    (per-pose reduction + combine -> loss and gradient, and the float64 set-up of the
    items for the next step); poses resident in HBM, loss and gradient left in HBM --
    and the fused kernel's mean duration from HIP events on its launch stream.  The
    kernel duration comes from KERNEL_SAMPLES further steps after the timed region, each
    with an event pair around its fused kernel (an event pair on every step adds about
    6 us to a 77 us step on MI355X, so the timed region itself carries a pair only on
    every KERNEL_EVENT_EVERY-th step; their mean is reported beside it)."""
    import torch
    ev.begin(P)
    for _ in range(warmup):
        ev.step(P)
    torch.cuda.synchronize()
    timed = range(0, steps, KERNEL_EVENT_EVERY)
    ks = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for k in timed}
    t0 = time.perf_counter()
    for k in range(steps):
        ev.step(P, kernel_events=ks.get(k))
    torch.cuda.synchronize()
    total_ms = (time.perf_counter() - t0) * 1e3
    in_region = [a.elapsed_time(b) for a, b in ks.values()]
    es = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(KERNEL_SAMPLES)]
    for e in es:
        ev.step(P, kernel_events=e)
    torch.cuda.synchronize()
    kern_ms = [a.elapsed_time(b) for a, b in es]
    return total_ms / steps, statistics.mean(kern_ms), kern_ms, statistics.mean(in_region)


def time_three(ev, P, steps: int, warmup: int):
    """Per-step time of the three-launch form (set-up, fused kernel, reduce + combine)."""
    import torch
    for _ in range(warmup):
        ev.evaluate_device(P, want_moments=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ev.evaluate_device(P, want_moments=False)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def time_graph(ev, P, steps: int, warmup: int):
    """Per-step time of the same two launches captured once as a HIP graph and replayed;
    returns ms per step."""
    import torch
    ev.begin(P)
    ev.step(P)  # sizes the workspace and the tile map
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ev.step(P)
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def torch_engine_t(reps: int = 30, warmup: int = 5):
    """The reference's own GPU benchmark workload for TorchDiffFVREngine.forward_spectral_ncc
    (experiments/exp29_gpu_triton_and_ablation.py:278-286, 391-393: 128^3 volume padded 2x,
    BF16 spectrum, out_res 256, B = 64 poses from randn, a random 256^2 target spectrum with
    the high-pass filter), on a seeded phantom: forward only and forward + backward."""
    import torch
    from fvr import TorchDiffFVREngine
    from fvr.phantom import ellipsoid_phantom
    dev = torch.device("cuda", 0)
    vol = ellipsoid_phantom(128, n_ellipsoids=12, seed=1234)
    eng = TorchDiffFVREngine(vol, pad_factor=2, device=dev)
    eng.to_precision(torch.bfloat16)
    g = torch.Generator(device="cpu").manual_seed(3456)
    B, H = 64, 256
    poses = torch.randn(B, 6, generator=g).to(dev)
    f_targ = eng.prepare_spectral_target(torch.randn(1, 1, H, H, generator=g).to(dev))

    def fwd():
        with torch.no_grad():
            return eng.forward_spectral_ncc(poses, f_targ)

    def fwd_bwd():
        p = poses.clone().requires_grad_(True)
        loss = eng.forward_spectral_ncc(p, f_targ)
        loss.backward()
        return loss

    out = {}
    for name, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[f"ms_{name}"] = round(ms, 4)
        out[f"pose_views_per_s_{name}"] = round(B / (ms * 1e-3), 1)
    return {"workload": "TorchDiffFVREngine.forward_spectral_ncc, 128^3 -> 256^3 bf16 spectrum, out_res 256, B=64, "
                        "high-pass, one view", **out,
            "reference_a100_fwd_pose_views_per_s": 49454.0}


def time_plain(ev, P, reps: int = 2):
    """One pose per launch sequence (set-up + kernel + reduce + combine per pose), FP32 spectrum."""
    import torch
    n = P.shape[0]
    singles = [P[k:k + 1].contiguous() for k in range(n)]
    for s in singles[:4]:
        ev.evaluate(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for s in singles:
            ev.evaluate(s)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / (reps * n)


def pou_config4(steps: int, warmup: int, size: int = 256, n_pose: int = 64):
    """PoU-FVR perspective workload: 256^3 seeded phantom, 4x4x4 blocks (25% overlap, cos^2
    windows, 64^3 resampled, P = 128 BF16 spectra), pinhole SDD 1020 mm / SAD 510 mm, 256^2
    detector at 0.8 mm, 64 candidate poses: forward (slice -> 128^2 LDS iFFT -> warp -> sum)
    + NCC loss vs a seeded target + adjoint to the 6-DoF pose parameters."""
    import numpy as np
    import torch
    from fvr import PerspectiveSlabFVR
    from fvr.phantom import ellipsoid_phantom
    dev = torch.device("cuda", 0)
    vol = ellipsoid_phantom(size, n_ellipsoids=12, seed=1234).astype(np.float32)
    t0 = time.perf_counter()
    model = PerspectiveSlabFVR.from_volume(vol, vox_mm=1.0, preset="isotropic", device=dev)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    det, pix, sdd, sad = 256, 0.8, 1020.0, 510.0
    f = sdd / pix
    k_inv = torch.tensor([[1.0 / f, 0.0, -det / (2 * f)], [0.0, 1.0 / f, -det / (2 * f)], [0.0, 0.0, 1.0]], device=dev)
    lattice = model.make_pinhole_lattice(k_inv, sdd, det, det)
    rng = np.random.default_rng(3456)

    def mats(rot_deg, trans_mm):
        from scipy.spatial.transform import Rotation
        M = np.tile(np.eye(4), (len(rot_deg), 1, 1))
        M[:, :3, :3] = Rotation.from_euler("zyx", rot_deg, degrees=True).as_matrix()
        M[:, :3, 3] = np.array([0.0, 0.0, -sad]) + trans_mm
        return torch.from_numpy(M).float().to(dev)

    true = mats(rng.uniform(-5, 5, (1, 3)), rng.uniform(-5, 5, (1, 3)))
    cands = mats(rng.uniform(-5, 5, (n_pose, 3)), rng.uniform(-5, 5, (n_pose, 3)))
    with torch.no_grad():
        target = model.forward_pinhole(true, lattice)
    from fvr import image_ncc

    def step():
        # NCC loss of the reference's PoU test (tests/test_cone_beam_fvr.py:24-27), per pose,
        # on the fused NCC kernels
        M = cands.clone().requires_grad_(True)
        out = model.forward_pinhole(M, lattice)
        loss = (1.0 - image_ncc(out, target)).sum()
        loss.backward()
        return loss, M.grad

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, g = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    with torch.no_grad():
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            model.forward_pinhole(cands, lattice)
        torch.cuda.synchronize()
        ms_fwd = (time.perf_counter() - t1) * 1e3 / steps
    # edge_dominant_ncc (cone_beam_fvr.py:71-112, the loss of the reference's PoU registration)
    # on the 64 projections against the target: forward, and forward + backward to the images
    from fvr import edge_dominant_ncc
    with torch.no_grad():
        proj = model.forward_pinhole(cands, lattice)
    pimg = proj.detach().clone().requires_grad_(True)
    for _ in range(warmup):
        edge_dominant_ncc(pimg, target).backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    with torch.no_grad():
        for _ in range(steps):
            edge_dominant_ncc(pimg, target)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(steps):
        pimg.grad = None
        edge_dominant_ncc(pimg, target).backward()
    torch.cuda.synchronize()
    e_fwd = (t3 - t2) * 1e3 / steps
    e_fb = (time.perf_counter() - t3) * 1e3 / steps
    img_b = n_pose * det * det * 4
    e_bytes = 3 * img_b + det * det * 4  # images read (forward, backward), gradient written; the target once
    edge = {"workload": f"edge_dominant_ncc, {n_pose} x {det}^2 projections vs one target (multiscale term on)",
            "ms_fwd": round(e_fwd, 4), "ms_fwd_bwd": round(e_fb, 4),
            "images_per_s_fwd_bwd": round(n_pose / (e_fb * 1e-3), 1),
            "algorithmic_bytes_fwd_bwd": e_bytes,
            "hbm_frac_fwd_bwd": round(e_bytes / (e_fb * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "grad_finite": bool(torch.isfinite(pimg.grad).all())}
    from fvr.footprint import pou_step_bytes, pou_texels_read
    R = cands[:, :3, :3].double().cpu().numpy()
    u_det = np.einsum("nij,j->ni", R, lattice[2].double().cpu().numpy())[:, ::-1]  # (x, y, z) storage order
    v_det = np.einsum("nij,j->ni", R, lattice[3].double().cpu().numpy())[:, ::-1]
    fp = pou_step_bytes(model.B, model.P, det * det, u_det, v_det, 4 if model.F_blk_ri.dtype != torch.float32 else 8,
                        texels_read=pou_texels_read(model, cands, lattice))
    gbs = fp["step"] / (ms * 1e-3) / 1e9
    return {"workload": "PoU-FVR 256^3, 4x4x4 blocks @64 (P=128, bf16), pinhole SDD 1020/SAD 510, 256^2 det @0.8mm",
            "poses": n_pose, "blocks": model.B, "P": model.P, "ms_fwd_adj": round(ms, 3), "ms_fwd": round(ms_fwd, 3),
            "pose_evals_per_s": round(n_pose / (ms * 1e-3), 1), "build_s": round(build_s, 2),
            "loss": float(loss.detach()), "grad_finite": bool(torch.isfinite(g).all()), "edge_ncc": edge,
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_step": fp["step"],
                         "bytes_by_kernel": {k: fp[k] for k in ("slice_ifft", "warp_fwd", "render_bwd")},
                         "unique_entries_per_pose_slice": fp["unique_entries_per_pose_mean"],
                         "slab_texels_read": fp["slab_texels_read"],
                         "note": "whole forward + adjoint step (slice + LDS iFFT, warp, fused adjoint) over 64 poses"}}


def pou_reference_workload(runs: int = 60, warmup: int = 15):
    """The PoU-FVR workloads of the reference's published GPU figures (exp29:259-270,
    296-311, 395-405; exp29_gpu_benchmark.csv:81, 92, 98-99, 103, 136): 2x2x2 blocks @ 64
    (P = 128), BF16 block spectra, pinhole detector of H^2 pixels at 2 * 128 / H mm, SDD
    1020 mm, source 850 mm from the isocentre, B poses with rotations (5, -3, 4) + N(0, 2^2)
    deg; forward + backward of forward_pinhole(...).sum() to the pose matrices, and the
    forward alone, at (B, H) = (64, 256) -- the headline point -- (16, 256), (1, 256) and
    (64, 512).  The reference ran them on a DeepFluoro CT, which is not available here: a
    seeded 256^3 ellipsoid phantom at 1 mm stands in (the cost follows the block count, P,
    the detector and the batch, not the volume's content)."""
    import numpy as np
    import torch
    from fvr import PerspectiveSlabFVR
    from fvr.phantom import ellipsoid_phantom
    from scipy.spatial.transform import Rotation
    dev = torch.device("cuda", 0)
    vol = ellipsoid_phantom(256, n_ellipsoids=12, seed=1234).astype(np.float32)
    model = PerspectiveSlabFVR.from_volume(vol, vox_mm=1.0, grid_blocks=(2, 2, 2), block_res=64, pad_mult=2,
                                           dtype=torch.bfloat16, device=dev)
    sdd, sid = 1020.0, 850.0

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(runs):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / runs, out

    def point(n_pose, det):
        pix = 2.0 * 128.0 / det
        f = sdd / pix
        k_inv = torch.tensor([[1.0 / f, 0.0, -det / (2 * f)], [0.0, 1.0 / f, -det / (2 * f)], [0.0, 0.0, 1.0]],
                             device=dev)
        lattice = model.make_pinhole_lattice(k_inv, sdd, det, det)
        rng = np.random.default_rng(29)
        rot = np.array([5.0, -3.0, 4.0]) + rng.normal(0.0, 2.0, (n_pose, 3))
        M = np.tile(np.eye(4), (n_pose, 1, 1))
        M[:, :3, :3] = Rotation.from_euler("zyx", rot, degrees=True).as_matrix()
        M[:, :3, 3] = (0.0, 0.0, -sid)
        poses = torch.from_numpy(M).float().to(dev)

        def fwd_bwd():
            rp = poses.clone().requires_grad_(True)
            model.forward_pinhole(rp, lattice).sum().backward()
            return rp.grad

        ms_fb, g = timed(fwd_bwd)
        with torch.no_grad():
            ms_f, _ = timed(lambda: model.forward_pinhole(poses, lattice))
        return {"B": n_pose, "det": det, "ms_fwd_bwd": round(ms_fb, 4), "ms_fwd": round(ms_f, 4),
                "poses_per_s_fwd_bwd": round(n_pose / (ms_fb * 1e-3), 1), "grad_finite": bool(torch.isfinite(g).all())}

    head = point(64, 256)
    ref = {(16, 256): 4.270, (1, 256): 4.210, (64, 512): 21.843}
    others = []
    for (b, h), ref_ms in ref.items():
        r = point(b, h)
        r["reference_a100_ms_fwd_bwd"] = ref_ms
        others.append(r)
    return {"workload": "PoU-FVR 2x2x2 blocks @64 (P=128, bf16), 256^2 pinhole det @1.0mm, SDD 1020 / source 850 mm, "
                        "B=64, forward_pinhole(...).sum() fwd+bwd (the reference's published PoU GPU workload; a 256^3 "
                        "phantom stands in for its CT)",
            "blocks": model.B, "P": model.P, "runs": runs, "ms_fwd_bwd": head["ms_fwd_bwd"], "ms_fwd": head["ms_fwd"],
            "poses_per_s_fwd_bwd": head["poses_per_s_fwd_bwd"],
            "poses_per_s_fwd": round(64 / (head["ms_fwd"] * 1e-3), 1), "grad_finite": head["grad_finite"],
            "reference_a100": {"ms_fwd_bwd": 7.074, "poses_per_s_fwd_bwd": 9047, "ms_fwd_eager": 2.703,
                               "ms_fwd_cuda_graph": 1.969},
            "other_points": others}


def reference_drr_workloads(reps: int = 10):
    """Two more of the reference's published timings, on this build's GPU path.
    (1) The convention-T `forward` DRR of exp24 (exp24_torch_gpu_benchmark.py:44-105,
    table24_torch_gpu_benchmark.csv:6): 96^3 volume, AP, out_res 96, B = 64 poses U(-10, 10)
    (degrees / voxels), forward per DRR and the backward of sum(drr^2); the reference's row
    has no device recorded (0.418 ms per DRR, backward 34.96 ms).  (2) The parallel-beam
    spectrum precompute and `render_drr` at 256^3 (exp1_complexity_scaling.py:33-49,
    table1_complexity_benchmark.csv:5: 6.38 s and 123.9 ms on the author's CPU).  Seeded
    ellipsoid phantoms stand in for the reference's clinical volume."""
    import numpy as np
    import torch
    from fvr import DifferentiableFourierVolumeRenderer, TorchDiffFVREngine
    from fvr.phantom import ellipsoid_phantom
    dev = torch.device("cuda", 0)
    out = {}
    eng = TorchDiffFVREngine(ellipsoid_phantom(96, n_ellipsoids=12, seed=1234).astype(np.float32),
                             spacing=(1.0, 1.0, 1.0), pad_factor=2.0, device=dev)
    rng = np.random.default_rng(24)
    poses = torch.tensor(rng.uniform(-10.0, 10.0, (64, 6)).astype(np.float32), device=dev, requires_grad=True)
    for _ in range(2):  # warm-up of both passes (FFT plans, autograd)
        torch.sum(eng(poses, view="ap", out_res=96) ** 2).backward()
        poses.grad = None
    torch.cuda.synchronize()
    fw, bw = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        drrs = eng(poses, view="ap", out_res=96)
        torch.cuda.synchronize()
        fw.append(time.perf_counter() - t0)
        loss = torch.sum(drrs ** 2)
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        bw.append(time.perf_counter() - t1)
        poses.grad = None
    fwd_ms = float(np.median(fw)) * 1e3
    out["torch_forward_drr_96"] = {"B": 64, "out_res": 96, "forward_ms": round(fwd_ms, 4),
                                   "per_drr_ms": round(fwd_ms / 64, 5), "backward_ms": round(float(np.median(bw)) * 1e3, 4),
                                   "timing": f"medians of {reps} calls after two warm-up forward + backward passes",
                                   "reference": {"per_drr_ms": 0.418, "backward_ms": 34.96, "device": "not recorded"}}
    vol = ellipsoid_phantom(256, n_ellipsoids=12, seed=1234).astype(np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = DifferentiableFourierVolumeRenderer(vol, pad_factor=2.0, device=dev)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    r.render_drr(0.0, 0.0, 0.0, 0.0, 0.0, "ap")
    ts = []
    for k in range(reps):
        t0 = time.perf_counter()
        img = r.render_drr(2.0 * k, -1.0, 3.0, 1.5, -2.0, "ap")
        ts.append(time.perf_counter() - t0)
    out["render_drr_256"] = {"prep_s": round(prep_s, 3), "render_drr_ms": round(float(np.median(ts)) * 1e3, 4),
                             "image": list(np.asarray(img).shape),
                             "note": "render_drr returns a host array: the time includes its device-to-host copy",
                             "reference_author_cpu": {"prep_s": 6.38, "render_drr_ms": 123.9}}
    return out


def multistart_config5(n_steps: int = 100, lat_reps: int = 100):
    """Ragged batched multistart: seeded ellipsoid phantoms of 128^3, 192^3 and 256^3
    (spectra 256^3 / 384^3 / 512^3, FP16 with one scale each), AP + LAT targets of
    128^2 / 192^2 / 256^2 DRRs at a seeded pose each, 512 starting poses split 96 / 160 /
    256 (the truth plus rotations U(-8, 8) deg, translations U(-8, 8) voxels).  All
    volumes go into ONE fused launch per step (autotuned launch plan); the 100-step Adam
    loop (reference hyper-parameters, tol 0) is captured once as a HIP graph and replayed.
    Reports pose evaluations per second of the replayed loop, the fused step's p50/p99
    latency (one-step graph replays timed with HIP events) and each volume's p50/p99 step
    latency when it runs alone, and the best-start mTRE per volume."""
    import numpy as np
    import torch
    from fvr import DifferentiableFourierVolumeRenderer
    from fvr.multistart import AdamLoop, autotune
    from fvr.phantom import ellipsoid_phantom, random_poses
    from fvr.registration_3d2d import compute_mtre_6dof
    from fvr.sncc import BiplanarEvaluator, RaggedBiplanarEvaluator
    dev = torch.device("cuda", 0)
    sizes, counts = (128, 192, 256), (96, 160, 256)
    segs, truths, starts = [], [], []
    t0 = time.perf_counter()
    for k, (n, c) in enumerate(zip(sizes, counts)):
        vol = ellipsoid_phantom(n, n_ellipsoids=12, seed=1234 + k)
        rng = np.random.default_rng(2345 + k)
        tp = np.concatenate([rng.uniform(-5, 5, 3), rng.uniform(-5, 5, 3)])
        r = DifferentiableFourierVolumeRenderer(vol, pad_factor=2.0, storage="fp16", device=dev)
        T_ap, T_lat = make_targets(r, tp)
        segs.append(BiplanarEvaluator(r.packed, T_ap, T_lat, adjoint=True))
        truths.append(tp)
        starts.append(tp + random_poses(c, 8.0, 8.0, seed=3456 + k))
        del vol
    build_s = time.perf_counter() - t0
    N = sum(counts)
    rag = RaggedBiplanarEvaluator(segs, counts)
    P0 = torch.from_numpy(np.ascontiguousarray(np.vstack(starts))).to(dev)
    plan, timings = autotune(rag, P0, reps=5)

    def lat_pcts(ev, n, P):
        one = AdamLoop(ev, n, 1, graph=True, tol=0.0)
        one.run(P)
        ev_s = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(lat_reps)]
        for a, b in ev_s:
            a.record()
            one.graph.replay()
            b.record()
        torch.cuda.synchronize()
        t = np.array([a.elapsed_time(b) for a, b in ev_s])
        return {"p50_ms": round(float(np.percentile(t, 50)), 4), "p99_ms": round(float(np.percentile(t, 99)), 4)}

    loop = AdamLoop(rag, N, n_steps, graph=True, tol=0.0)
    loop.run(P0)  # capture + first run
    torch.cuda.synchronize()
    reps = 3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_host = time.perf_counter()
    e0.record()
    for _ in range(reps):
        best, best_loss, _ = loop.run(P0)
    e1.record()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t_host) * 1e3 / reps
    loop_ms = e0.elapsed_time(e1) / reps
    fused = lat_pcts(rag, N, P0)
    per_volume, mtre = {}, {}
    best, best_loss = best.cpu().numpy(), best_loss.cpu().numpy()
    off = 0
    for k, (n, c) in enumerate(zip(sizes, counts)):
        per_volume[f"{n}^3"] = dict(poses=c, **lat_pcts(segs[k], c, P0[off:off + c].contiguous()))
        kb = off + int(np.argmin(best_loss[off:off + c]))
        mtre[f"{n}^3"] = round(float(compute_mtre_6dof(truths[k], best[kb], (n, n, n))), 4)
        off += c
    return {"workload": "ragged multistart: 128^3/192^3/256^3 phantoms (256^3/384^3/512^3 fp16 spectra), "
                        "AP+LAT targets, 512 starts split 96/160/256, one fused launch per step, "
                        f"{n_steps}-step Adam captured as one HIP graph",
            "poses": N, "steps": n_steps, "loop_ms": round(loop_ms, 3), "loop_wall_ms": round(wall_ms, 3),
            "pose_evals_per_s": round(N * n_steps / (loop_ms * 1e-3), 1),
            "fused_step_latency": fused, "per_volume_step_latency_alone": per_volume,
            "launch_plan": {"tile_h": plan.tile_h, "order": plan.order, "persist": plan.persist,
                            "depth": plan.depth},
            "autotune_ms": {f"{p.tile_h}/{p.order}/{p.persist}/{p.depth}": round(t, 4) for p, t in
                            sorted(timings.items(), key=lambda kv: kv[1])},
            "best_start_mtre_mm": mtre, "build_s": round(build_s, 2)}


_POOL_SPEC = None


def _pool_init(path: str, ta, tl) -> None:
    """Worker set-up of the all-cores CPU baseline: map the spectrum read-only."""
    global _POOL_SPEC
    import numpy as np
    _POOL_SPEC = (np.load(path, mmap_mode="r"), ta, tl)


def _pool_eval(pose):
    from oracle import diff_fvr_np as O
    f3, ta, tl = _POOL_SPEC
    t = time.perf_counter()
    O.biplanar_loss_and_gradients(f3, pose, ta, tl)
    return time.perf_counter() - t


def host_cores() -> int:
    """Worker processes for the all-cores CPU baseline: the cores this process may run on,
    at most 16 (a GPU box's host share per GPU)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def powell_registration(size: int = 128):
    """The reference's Powell registration driver (core_engine.py:109-224, clean Spectral-NCC
    objective) on a seeded 128^3 phantom at 1 mm: DRR targets at a seeded pose, start
    offset by (4, -3, 2.5 deg, 5, -4, 3 mm); scipy's Powell (one pose per fused loss-only
    launch, as the reference asks for them) against the batched Powell (a grid of poses
    along each line per launch).  Evaluations, launches, wall time, evaluations per
    second and mTRE."""
    import numpy as np
    from fvr import RobustDiffFVREngine
    from fvr.phantom import ellipsoid_phantom
    from fvr.registration_3d2d import compute_mtre_6dof
    vol = ellipsoid_phantom(size, n_ellipsoids=12, seed=1234).astype(np.float64)
    eng = RobustDiffFVREngine(vol, spacing=(1.0,) * 3, pad_factor=2.0)
    rng = np.random.default_rng(2345)
    truth = np.concatenate([rng.uniform(-5, 5, 3), rng.uniform(-5, 5, 3)])
    ap = eng.render_calibrated_drr(*truth[:3], tx=truth[3], ty=truth[5], view="ap")
    lat = eng.render_calibrated_drr(*truth[:3], tx=truth[4], ty=truth[5], view="lateral")
    init = truth + np.array([4.0, -3.0, 2.5, 5.0, -4.0, 3.0])
    out = {"workload": f"RobustDiffFVREngine.register_biplanar_robust, {size}^3 phantom at 1 mm, "
                       f"{2 * size}^3 spectrum, clean Spectral-NCC objective, max_iter 50"}
    for opt in ("powell", "batched_powell"):
        eng.register_biplanar_robust(ap, lat, initial_pose=tuple(init), max_iter=2, optimizer=opt)  # warm-up
        res = eng.register_biplanar_robust(ap, lat, initial_pose=tuple(init), max_iter=50, optimizer=opt)
        m = compute_mtre_6dof(tuple(truth), tuple(res["final_pose"]), vol.shape, translation_units="mm")
        out[opt] = {"evaluations": res["evaluations"], "launches": res["launches"],
                    "time_s": round(res["time_sec"], 4),
                    "evals_per_s": round(res["evaluations"] / res["time_sec"], 1),
                    "final_loss": res["final_loss"], "mtre_mm": round(float(m), 4)}
    out["speedup_wall"] = round(out["powell"]["time_s"] / out["batched_powell"]["time_s"], 2)
    return out


def cpu_baseline(renderer, T_ap, T_lat, cands, budget_s: float, gpu=None):
    """The NumPy/SciPy oracle (reference algorithm: 14 map_coordinates slices per pose
    eval) on bounded samples of the candidate poses: (a) one host core, in this process;
    (b) a process pool over independent candidates on every host core this process may
    use (the spectrum shared read-only through a memory-mapped file).  ``gpu``: {storage:
    (loss (N,), grad (N, 6))} of the GPU launch -- the sampled oracle losses, and the exact
    oracle gradient of two of them, are compared with it (max |dloss|, max |dgrad| / max|grad|)."""
    import multiprocessing as mp
    import tempfile

    import numpy as np
    from oracle import diff_fvr_np as O
    f3 = renderer.f_3d
    ta, tl = T_ap.cpu().numpy().astype(np.complex128), T_lat.cpu().numpy().astype(np.complex128)
    spec_shape = "x".join(str(n) for n in f3.shape)
    times, ref = [], {}
    t_start = time.perf_counter()
    k = 0
    while k < len(cands) and (k < 1 or time.perf_counter() - t_start < budget_s):
        t = time.perf_counter()
        ref[k] = O.biplanar_loss_and_gradients(f3, cands[k], ta, tl)
        times.append(time.perf_counter() - t)
        k += 1
        log(f"cpu baseline (1 core): pose {k} took {times[-1]:.2f} s")
    med = statistics.median(times)
    parity = {}
    if gpu:
        exact = {j: O.biplanar_loss_and_grad_analytic(f3, cands[j], ta, tl, fold=True) for j in (0, k - 1)}
        for st, (L, G) in gpu.items():
            dl = max(abs(float(L[j]) - ref[j][0]) for j in ref)
            dg = max(float(np.abs(G[j] - g).max() / np.abs(g).max()) for j, (_, g) in exact.items())
            parity[st] = {"max_abs_dloss": dl, "max_rel_dgrad": dg}
        parity["candidates_loss"] = sorted(ref)
        parity["candidates_grad"] = sorted(exact)
        parity["oracle"] = ("loss: the oracle's reference scheme; gradient: the oracle's exact adjoint "
                            "(biplanar_loss_and_grad_analytic), both FP64 on the same spectrum and targets")
        log(f"oracle parity at full size: {parity}")
    # (b) every host core: a process pool over candidate poses
    n_work = host_cores()
    tmp = tempfile.mkdtemp(dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    path = os.path.join(tmp, "spectrum.npy")
    pool_value, pool_sample = None, None
    try:
        np.save(path, f3)
        del f3
        old = os.environ.get("OMP_NUM_THREADS")
        os.environ["OMP_NUM_THREADS"] = "1"
        try:
            ctx = mp.get_context("spawn")
            with ctx.Pool(n_work, initializer=_pool_init, initargs=(path, ta, tl)) as pool:
                pool.map(_pool_eval, [cands[j % len(cands)] for j in range(n_work)], chunksize=1)  # warm-up
                n_eval = 0
                t0 = time.perf_counter()
                while n_eval == 0 or time.perf_counter() - t0 < budget_s:
                    batch = [cands[(n_eval + j) % len(cands)] for j in range(n_work)]
                    pool.map(_pool_eval, batch, chunksize=1)
                    n_eval += len(batch)
                wall = time.perf_counter() - t0
            pool_value = n_eval / wall
            pool_sample = f"{n_eval} pose evals in {wall:.1f} s of wall time on {n_work} worker processes"
            log(f"cpu baseline ({n_work} cores): {pool_value:.2f} pose evals/s")
        finally:
            if old is None:
                os.environ.pop("OMP_NUM_THREADS", None)
            else:
                os.environ["OMP_NUM_THREADS"] = old
    finally:
        if os.path.exists(path):
            os.remove(path)
        os.rmdir(tmp)
    one = {"value": 1.0 / med, "cores": 1,
           "sample": f"{k} of the {len(cands)} candidate poses, median of {k} evals"}
    base = (f"candidate poses on the {spec_shape} complex128 spectrum, reference FD scheme (14 scipy "
            "map_coordinates slices per pose eval)")
    out = {"value": pool_value, "unit": "pose evals/s", "cores": n_work, "kind": "port",
           "sample": f"{pool_sample}; {base}", "one_core": one, "host_cpu_count": os.cpu_count(),
           "host_cores_available": len(os.sched_getaffinity(0))}
    return out, parity


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--n-cand", type=int, default=64)
    ap.add_argument("--storage", default="fp16", choices=["f32", "fp16", "bf16"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="headline storage only (no f32 / plain forms)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-pou", action="store_true", help="skip the PoU-FVR perspective workload")
    ap.add_argument("--no-multistart", action="store_true", help="skip the ragged multistart workload")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay timing")
    ap.add_argument("--plan", default="auto",
                    help="fused-kernel launch plan 'tile_h/order/persist[/depth]' (e.g. 64/target/0/3) or 'auto' "
                         "(autotune)")
    ap.add_argument("--no-footprint", action="store_true",
                    help="skip the distinct-entry count (roofline fields null; for profiler runs)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "candidate poses would shard across GPUs with the spectrum replicated",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from fvr import DifferentiableFourierVolumeRenderer
    from fvr.footprint import batch_algorithmic_bytes
    from fvr.multistart import autotune
    from fvr.sncc import BiplanarEvaluator, LaunchPlan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log(f"device {torch.cuda.get_device_name(dev)}; building {args.size}^3 phantom")
    vol, true_pose, cands = make_inputs(args.size, args.n_cand)
    log("spectrum precompute (float64, device) + packing")
    r = DifferentiableFourierVolumeRenderer(vol, pad_factor=2.0, storage="f32", device=dev)
    T_ap, T_lat = make_targets(r, true_pose)
    P = torch.from_numpy(np.ascontiguousarray(cands, dtype=np.float64)).to(dev)
    results = {}
    storages = [args.storage] if args.no_extra else sorted({args.storage, "f32"}, key=lambda s: s != args.storage)
    for st in storages:
        if r.storage != st:
            r.repack(st)
        ev = BiplanarEvaluator(r.packed, T_ap, T_lat, adjoint=True)
        timings = {}
        if args.plan == "auto":
            plan, timings = autotune(ev, P, reps=5)
        else:
            th, order, persist, *depth = args.plan.split("/")
            plan = LaunchPlan(int(th), order, int(persist), int(depth[0]) if depth else 2)
            ev.set_plan(plan)
        step_ms, kern_ms, kern_all, kern_region = time_batched(ev, P, args.steps, args.warmup)
        graph_ms = time_graph(ev, P, args.steps, args.warmup) if not args.no_graph else None
        three_ms = time_three(ev, P, args.steps, args.warmup)
        loss, grad = ev.evaluate(cands)
        # NumPy poses in, NumPy loss / gradient / moments out: the host-buffer boundary (PCIe copies included)
        torch.cuda.synchronize()
        t_h = time.perf_counter()
        for _ in range(10):
            ev.evaluate(cands)
        host_ms = (time.perf_counter() - t_h) * 1e3 / 10
        c_abi = None
        if st == args.storage and not args.no_extra:
            # the one-call C entry point (fvr_pn_evaluate) with the plan its own C autotuner
            # picks (fvr_autotune_pn_evaluate): what a caller outside Python gets per batch
            c_plan, c_tuned_ms = ev.autotune_call(P, reps=5)
            for _ in range(args.warmup):
                ev.evaluate_call(P)
            torch.cuda.synchronize()
            t_c = time.perf_counter()
            for _ in range(args.steps):
                ev.evaluate_call(P)
            torch.cuda.synchronize()
            c_ms = (time.perf_counter() - t_c) * 1e3 / args.steps
            c_abi = {"plan": f"{c_plan.tile_h}/{c_plan.order}/{c_plan.depth}", "autotune_ms": round(c_tuned_ms, 4),
                     "ms_per_call": round(c_ms, 4), "pose_evals_per_s": round(args.n_cand / (c_ms * 1e-3), 1)}
        results[st] = {"c_abi": c_abi, "ms_per_step": step_ms, "kernel_ms": kern_ms, "kernel_ms_samples": len(kern_all),
                       "kernel_ms_p50": statistics.median(kern_all), "kernel_ms_min": min(kern_all),
                       "kernel_ms_in_timed_region": kern_region, "value": args.n_cand / (step_ms * 1e-3),
                       "loss_mean": float(loss.mean()), "loss": loss, "grad": grad, "herm": ev.herm,
                       "plan": plan, "autotune": timings, "host_numpy_ms_per_step": host_ms,
                       "three_launch_ms_per_step": three_ms, "graph_ms_per_step": graph_ms}
        log(f"{st}: plan {plan}: {step_ms:.3f} ms/step, fused kernel {kern_ms:.3f} ms, "
            f"{results[st]['value']:.0f} pose evals/s")
        del ev
    plain_ms = None
    if not args.no_extra:
        if r.storage != "f32":
            r.repack("f32")
        ev = BiplanarEvaluator(r.packed, T_ap, T_lat, adjoint=True, herm=False)  # plain: full slice, one pose
        plain_ms = time_plain(ev, P)
        log(f"plain one-pose-per-launch f32: {plain_ms:.3f} ms/pose")
        del ev
    head = results[args.storage]
    elem = 8 if args.storage == "f32" else 4
    if args.no_footprint:
        fp = {"bytes": None, "unique_entries": None,
              "schedule": "hermitian-fold" if head["herm"] else "full-slice"}
        achieved = None
    else:
        log("counting the algorithmic footprint (distinct spectrum entries per slice)")
        fp = batch_algorithmic_bytes((2 * args.size,) * 3, cands, elem, scaled=args.storage != "f32", device=dev,
                                     herm=head["herm"])
        achieved = fp["bytes"] / (head["kernel_ms"] * 1e-3) / 1e9
    traffic = None
    kname = "sncc_herm_kernel" if head["herm"] else "sncc_kernel"
    pmc_path = os.path.join(ROOT, "profiles", f"{kname}_{args.storage}_{args.size}cube_{args.n_cand}cand.json")
    if os.path.exists(pmc_path):
        pmc = json.load(open(pmc_path))
        traffic = {"bytes": pmc["traffic_bytes_per_launch"], "l2_hit_rate": pmc["l2_hit_rate"],
                   "source": os.path.relpath(pmc_path, ROOT), "pmc_avg_duration_us": pmc["avg_duration_us"]}
    cpu = None
    extra = {}
    if not args.no_cpu:
        log("CPU baseline (oracle port: one core, then every host core) + oracle parity of the sampled poses")
        cpu, extra["oracle_parity"] = cpu_baseline(r, T_ap, T_lat, cands, args.cpu_seconds,
                                                   {st: (v["loss"], v["grad"]) for st, v in results.items()})
    if "f32" in results and args.storage != "f32":
        extra["f32_batched_value"] = results["f32"]["value"]
        extra["f32_batched_ms_per_step"] = results["f32"]["ms_per_step"]
        extra["f32_vs_storage_max_abs_dloss"] = float(np.max(np.abs(results["f32"]["loss"] - head["loss"])))
        g32 = results["f32"]["grad"]
        extra["f32_vs_storage_max_rel_dgrad"] = float(np.max(np.abs(g32 - head["grad"])) / np.max(np.abs(g32)))
    if not args.no_extra and not args.no_pou:
        log("PoU-FVR perspective workload (config 4)")
        del r
        torch.cuda.empty_cache()
        extra["pou"] = pou_config4(max(3, args.steps // 10), 2)
        log(f"PoU: {extra['pou']}")
        extra["pou_reference_workload"] = pou_reference_workload()
        log(f"PoU (the reference's published workload): {extra['pou_reference_workload']}")
        extra["reference_drr_workloads"] = reference_drr_workloads()
        log(f"DRR workloads of the reference's published timings: {extra['reference_drr_workloads']}")
    if not args.no_extra:
        log("convention-T engine on the reference's own GPU benchmark workload")
        torch.cuda.empty_cache()
        extra["torch_engine_t"] = torch_engine_t()
        log(f"torch engine: {extra['torch_engine_t']}")
    if not args.no_extra:
        log("Powell registration drivers (scipy vs batched line searches)")
        torch.cuda.empty_cache()
        extra["powell"] = powell_registration()
        log(f"powell: {extra['powell']}")
    if not args.no_extra and not args.no_multistart:
        log("ragged multistart workload (config 5)")
        torch.cuda.empty_cache()
        extra["multistart"] = multistart_config5()
        log(f"multistart: {extra['multistart']}")
    extra["three_launch_ms_per_step"] = round(head["three_launch_ms_per_step"], 4)
    if head["c_abi"] is not None:
        extra["c_abi_one_call"] = head["c_abi"]
    if head["graph_ms_per_step"] is not None:
        extra["graph_replay_ms_per_step"] = round(head["graph_ms_per_step"], 4)
    extra["host_numpy_ms_per_step"] = round(head["host_numpy_ms_per_step"], 4)
    extra["host_numpy_pose_evals_per_s"] = round(args.n_cand / (head["host_numpy_ms_per_step"] * 1e-3), 1)
    if plain_ms is not None:
        extra["plain_f32_per_pose_value"] = 1e3 / plain_ms
        extra["speedup_vs_plain_f32"] = head["value"] / (1e3 / plain_ms)
    out = {
        "metric": METRIC,
        "value": round(head["value"], 1),
        "unit": "pose evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": args.storage,
        "data": "synthetic",
        "config": {"workload": "parallel-beam biplanar Spectral-NCC fwd + exact 6-DoF adjoint",
                   "volume": args.size, "spectrum": 2 * args.size, "views": ["ap", "lateral"],
                   "slice_samples_per_view": (2 * args.size) ** 2, "candidates_per_launch": args.n_cand,
                   "spectrum_storage": args.storage,
                   "launch_plan": {"tile_h": head["plan"].tile_h, "order": head["plan"].order,
                                   "persist": head["plan"].persist, "depth": head["plan"].depth,
                                   "autotuned": args.plan == "auto"}},
        "roofline": {"bound": "hbm", "achieved": None if achieved is None else round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else traffic["bytes"], "traffic_detail": traffic,
                     "kernel": kname, "kernel_ms": round(head["kernel_ms"], 4),
                     "kernel_ms_samples": head["kernel_ms_samples"], "kernel_ms_p50": round(head["kernel_ms_p50"], 4),
                     "kernel_ms_min": round(head["kernel_ms_min"], 4),
                     "kernel_ms_in_timed_region": round(head["kernel_ms_in_timed_region"], 4),
                     "algorithmic_bytes_per_launch": fp["bytes"], "unique_entries": fp["unique_entries"],
                     "schedule": fp.get("schedule"), "samples_per_launch": fp.get("samples")},
        "cpu_baseline": cpu,
        "autotune_ms": {f"{p.tile_h}/{p.order}/{p.persist}/{p.depth}": round(t, 4)
                        for p, t in sorted(head["autotune"].items(), key=lambda kv: kv[1])},
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
