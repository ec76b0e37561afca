"""ctypes binding of ``libfvr_hip.so`` (the C ABI declared in ``include/fvr_hip.h``).

The library links the HIP runtime torch ships, so ``torch`` is imported first
and the library is then opened once per process.  There is no fallback: if
the library is missing, every product entry point raises.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import torch  # noqa: F401  (must be loaded before the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfvr_hip.so")
ABI_VERSION = 4

FVR_C32, FVR_C16F, FVR_C16B = 0, 1, 2
DTYPE_CODES = {"f32": FVR_C32, "fp16": FVR_C16F, "bf16": FVR_C16B}
ITEM_ADJOINT, ITEM_HP, ITEM_WRITE_SLICE, ITEM_HERM = 1, 2, 4, 8
LAYOUT_WRAP = 1
NMOM = 16

# numpy mirrors of the C structs (explicit offsets; sizes are checked against
# the library at load time)
LAYOUT_DTYPE = np.dtype({
    "names": ["base", "scales", "n", "dtype", "sdim", "reserved", "elem_bytes", "row_bytes", "plane_bytes",
              "flags", "total_bytes"],
    "formats": ["<u8", "<u8", ("<i4", 3), "<i4", ("<i4", 3), ("<i4", 3), "<u4", "<u4", "<u4", "<u4", "<u8"],
    "offsets": [0, 8, 16, 28, 32, 44, 56, 60, 64, 68, 72],
    "itemsize": 80,
})

ITEM_DTYPE = np.dtype({
    "names": ["c", "ustep", "vstep", "box_hi", "ph_col0", "ph_col", "ph_row0", "ph_row", "cw0", "cws", "rw0",
              "rws", "hp_inv2s2", "hp_x0", "hp_dx", "hp_y0", "hp_dy", "h", "w", "layout", "lo_int", "guard_off",
              "tile_base", "tiles_x", "flags", "tile_h", "target", "slice_out"],
    "formats": [("<i8", 3), ("<i8", 3), ("<i8", 3), ("<u8", 3), "<u4", "<u4", "<u4", "<u4", "<f4", "<f4", "<f4",
                "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<i4", "<i4", "<i4", "<i4", "<u4", "<i4", "<i4", "<u4",
                "<i4", "<u8", "<u8"],
    "offsets": [0, 24, 48, 72, 96, 100, 104, 108, 112, 116, 120, 124, 128, 132, 136, 140, 144, 148, 152, 156, 160,
                164, 168, 172, 176, 180, 184, 192],
    "itemsize": 200,
})

ADAM_DTYPE = np.dtype({
    "names": ["lr", "gclip", "lo", "hi", "beta1", "beta2", "eps", "tol", "cosine_t", "pad0"],
    "formats": [("<f8", 6), ("<f8", 6), ("<f8", 6), ("<f8", 6), "<f8", "<f8", "<f8", "<f8", "<i4", "<i4"],
    "offsets": [0, 48, 96, 144, 192, 200, 208, 216, 224, 228],
    "itemsize": 232,
})


class FvrNativeError(RuntimeError):
    """Raised when a libfvr_hip entry point returns a non-zero status."""


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int32
_U = ctypes.c_uint32
_U64 = ctypes.c_uint64
_F = ctypes.c_float

_SIGNATURES = {
    "fvr_abi_version": ([], _I),
    "fvr_last_error": ([], ctypes.c_char_p),
    "fvr_struct_sizes": ([_P], _I),
    "fvr_layout_init": ([_P, _P, _I, _P, _P], _I),
    "fvr_spec_pack": ([_P, _U64, _I, _I, _I, _I, _P], _I),
    "fvr_sncc_tile_rows": ([], _I),
    "fvr_sncc_has_tile_h": ([_U, _I], _I),
    "fvr_sncc_has_depth": ([_U, _I, _I], _I),
    "fvr_sncc_eval_v": ([_P, _I, _P, _I, _I, _P, _U, _P], _I),
    "fvr_sncc_eval_plan": ([_P, _I, _P, _I, _I, _P, _U, _P, _I, _I, _P], _I),
    "fvr_sncc_reduce": ([_P, _I, _P, _P, _P], _I),
    "fvr_pn_build_items": ([_P, _I, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I, _I, _P, _P, _P],
                           _I),
    "fvr_adam_step": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P], _I),
    "fvr_pt_build_items": ([_P, _I, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I, _I,
                            _P, _U64, _U64, _P, _P, _P], _I),
    "fvr_pt_combine": ([_P, _I, _P, _P, _P, _I, ctypes.c_double, _P, _P, _P], _I),
    "fvr_mean_loss": ([_P, _I, _P, _P], _I),
    "fvr_pn_step": ([_P, _I, _P, _P, ctypes.c_double, ctypes.c_double, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                     _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I, _I, _P], _I),
    "fvr_pn_reduce_combine": ([_P, _I, _P, _P, ctypes.c_double, ctypes.c_double, _P, _P, _P, _P], _I),
    "fvr_pn_evaluate_workspace": ([_P, _I], ctypes.c_int64),
    "fvr_pn_evaluate": ([_P, _I, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I, _U, _P, _I,
                         ctypes.c_double, ctypes.c_double, _P, ctypes.c_int64, _P, _P, _P, _P], _I),
    "fvr_pn_evaluate_noalloc": ([_P, _I, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I, _U, _P, _I,
                                 ctypes.c_double, ctypes.c_double, _P, _P, _P, _P], _I),
    "fvr_autotune_pn_evaluate": ([_P, _I, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I, _U, _P, _I,
                                  ctypes.c_double, ctypes.c_double, _P, ctypes.c_int64, _P, _P, _I, _I, _P, _P, _P],
                                 _I),
    "fvr_autotune_clear": ([], _I),
    "fvr_c64_pair_moments": ([_U64, _U64, _I, _I, _P, _P], _I),
    "fvr_slice_fwd": ([_U64, _I, _I, _I, _I, _I, _P, _P, _I, _P, _P, _I, _F, _F, _F, _U64, _P], _I),
    "fvr_slice_bwd_workspace": ([_I, _I], ctypes.c_int64),
    "fvr_slice_bwd": ([_U64, _I, _I, _I, _I, _I, _U64, _P, _P, _I, _P, _P, _I, _F, _F, _F, _P, _P, _P, _P], _I),
    "fvr_pou_lds_ok": ([_I], _I),
    "fvr_pou_set_mirror_queue": ([_I], _I),
    "fvr_pou_packed_bytes": ([_I, _I, _I], ctypes.c_int64),
    "fvr_pou_pack_blocks": ([_U64, _I, _I, _I, _U64, _P], _I),
    "fvr_pou_slice_ifft": ([_U64, _I, _I, _I, _P, _P, _I, _P, _P], _I),
    "fvr_pou_warp_fwd": ([_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _P, _P], _I),
    "fvr_pou_warp_fwd_sep": ([_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _P, _P], _I),
    "fvr_pou_gscale": ([_P, _P, _I, _I, _I, _P, _P], _I),
    "fvr_pou_scale_of_max": ([_F], _F),
    "fvr_pou_warp_bwd_workspace": ([_I, _I, _I], ctypes.c_int64),
    "fvr_pou_warp_bwd": ([_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _P, _P, _P, _P, _P], _I),
    "fvr_pou_render_bwd": ([_U64, _I, _I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _P, _P,
                            _P], _I),
    "fvr_pou_sep_ok": ([_I, _I, _I], _I),
    "fvr_pou_gprep": ([_P, _P, _I, _I, _I, _P, _P, _P], _I),
    "fvr_pou_render_bwd_sep": ([_U64, _I, _I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _I,
                                _P, _P], _I),
    "fvr_pou_ray_grads": ([_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _P, _P, _P, _P], _I),
    "fvr_pou_pose_geom": ([_P, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P], _I),
    "fvr_pou_pose_geom_bwd": ([_P, _I, _P, _P, _I, _P, _P, _P], _I),
    "fvr_pou_sum_blocks": ([_P, _I, _I, _I, _I, _I, _P, _P], _I),
    "fvr_edge_tiles": ([_I, _I, _I], _I),
    "fvr_edge_channels": ([_P, _I, _I, _I, _I, _I, _P, _P], _I),
    "fvr_edge_stats": ([_P, _I, _I, _I, _I, _I, _P, _I, _P, _P], _I),
    "fvr_edge_combine": ([_P, _I, _P, _I, _I, _I, _I, _I, ctypes.c_double, _P, _P, _P], _I),
    "fvr_edge_bwd": ([_P, _I, _I, _I, _I, _I, _P, _I, _P, _P, _P, _P], _I),
    "fvr_ncc_work_doubles": ([_I, ctypes.c_int64], ctypes.c_int64),
    "fvr_ncc_fwd": ([_P, _I, _P, _I, ctypes.c_int64, _P, _P, _P, _P], _I),
    "fvr_ncc_bwd": ([_P, _I, _P, _I, ctypes.c_int64, _P, _P, _I, _P, _P], _I),
}


def lib():
    """Open the library (once) and bind the ABI; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libfvr_hip.so not found at {LIB_PATH}; build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or `make -C diff-fvr_amd/csrc`).  There is no CPU fallback.")
    handle = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.argtypes = args
        fn.restype = res
    ver = handle.fvr_abi_version()
    if ver != ABI_VERSION:
        raise ImportError(f"libfvr_hip.so ABI {ver} != expected {ABI_VERSION}; rebuild the library")
    sizes = (ctypes.c_int32 * 4)()
    handle.fvr_struct_sizes(sizes)
    if (sizes[0], sizes[1], sizes[2]) != (LAYOUT_DTYPE.itemsize, ITEM_DTYPE.itemsize, ADAM_DTYPE.itemsize):
        raise ImportError(f"struct size mismatch: C ({sizes[0]}, {sizes[1]}, {sizes[2]}) vs Python "
                          f"({LAYOUT_DTYPE.itemsize}, {ITEM_DTYPE.itemsize}, {ADAM_DTYPE.itemsize})")
    _lib = handle
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().fvr_last_error().decode(errors="replace")
        raise FvrNativeError(f"{what} failed (status {status}): {msg}")


def require_device(t: torch.Tensor, what: str) -> None:
    """The HIP path only: CPU tensors are an error, never a silent fallback."""
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise RuntimeError(f"{what}: expects a tensor on a ROCm device (got "
                           f"{'CPU tensor' if isinstance(t, torch.Tensor) else type(t).__name__}); "
                           "this package has no CPU path")


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("fvr needs a ROCm GPU (torch.cuda.is_available() is False); there is no CPU path")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(device: torch.device | None = None) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def as_struct(arr, dtype: np.dtype) -> np.ndarray:
    """``arr`` as a contiguous array of exactly ``dtype`` (the C layout).

    NumPy operations such as ``np.concatenate`` return structured arrays with a
    *packed* dtype (alignment holes dropped, later fields shifted); converting
    by field name restores the C offsets and item size.
    """
    a = np.asarray(arr)
    if a.dtype != dtype or a.dtype.itemsize != dtype.itemsize:
        a = a.astype(dtype)
    return np.ascontiguousarray(a)


def struct_bytes(arr, dtype: np.dtype) -> np.ndarray:
    """Raw bytes of ``arr`` in the C layout of ``dtype`` (size checked)."""
    a = as_struct(arr, dtype)
    raw = a.view(np.uint8).reshape(-1)
    if raw.size != a.size * dtype.itemsize:
        raise AssertionError("struct byte image does not match the C item size")
    return raw


def struct_to_device(arr: np.ndarray, device: torch.device, dtype: np.dtype = None) -> torch.Tensor:
    """Upload a numpy structured array in its C layout as raw bytes (uint8 tensor)."""
    raw = struct_bytes(arr, dtype if dtype is not None else np.asarray(arr).dtype)
    return torch.from_numpy(raw.copy()).to(device, non_blocking=False)
