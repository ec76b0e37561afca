"""CPU oracle for the Diff-FVR rendering hot path.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or as the timed CPU
baseline), never as a fallback for the HIP path.

Modules
-------
``diff_fvr_np``      FP64 NumPy/SciPy restatement of the parallel-beam path
                     (convention N: ``src/fvr/diff_fvr.py``, ``gradients.py``,
                     ``gradient_optimizer.py``, ``core_engine.py``,
                     ``registration_3d2d.py``) plus the exact analytic pose
                     adjoint the HIP kernel implements.
``torch_engine_np``  FP64 NumPy restatement of convention T
                     (``src/fvr/torch_engine.py``: geometry, slice, image-space
                     loss) and of the generic per-corner slice operator of
                     ``src/fvr/triton_slice.py``.
``cone_beam_np``     FP64 NumPy restatement of the PoU-FVR pinhole path
                     (``src/fvr/cone_beam_fvr.py``: blocks, slab images, warp,
                     pinhole / C-arm projections, ``edge_dominant_ncc``).

How each function is pinned (golden vectors, known answers, closed forms,
brute force) is listed in each module's header; functions without such a
check say "parity unpinned" there.
"""
