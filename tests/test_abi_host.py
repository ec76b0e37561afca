"""CPU checks of the C ABI library and of the host-side logic (no GPU compute)."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fvr_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvr_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    from fvr import _lib
    L = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.fvr_abi_version() == _lib.ABI_VERSION


def test_python_bindings_cover_the_header():
    from fvr import _lib
    assert set(_declared_functions()) <= set(_lib._SIGNATURES), \
        set(_declared_functions()) - set(_lib._SIGNATURES)


def test_struct_sizes_match_numpy_mirrors():
    from fvr import _lib
    L = _lib.lib()
    out = (ctypes.c_int32 * 4)()
    assert L.fvr_struct_sizes(out) == 0
    assert out[0] == _lib.LAYOUT_DTYPE.itemsize == 80
    assert out[1] == _lib.ITEM_DTYPE.itemsize == 200
    assert out[2] == _lib.ADAM_DTYPE.itemsize == 232


def test_layout_init_and_error_reporting():
    from fvr import _lib
    L = _lib.lib()
    lay = np.zeros(1, dtype=_lib.LAYOUT_DTYPE)
    nb, ns = ctypes.c_uint64(), ctypes.c_uint64()
    n = np.array([512, 512, 512], dtype=np.int32)
    assert L.fvr_layout_init(lay.ctypes.data, n.ctypes.data, 0, ctypes.byref(nb), ctypes.byref(ns)) == 0
    assert tuple(lay["sdim"][0]) == (515, 514, 514)
    assert lay["row_bytes"][0] == 514 * 8 and lay["plane_bytes"][0] == 514 * 514 * 8
    assert nb.value == 515 * 514 * 514 * 8 and ns.value == 0
    assert L.fvr_layout_init(lay.ctypes.data, n.ctypes.data, 1, ctypes.byref(nb), ctypes.byref(ns)) == 0
    # scaled storage: the same storage dims with 4-byte elements and one scale per spectrum
    assert ns.value == 1 and lay["elem_bytes"][0] == 4 and lay["flags"][0] == 0
    assert tuple(lay["reserved"][0]) == (0, 0, 0)
    assert lay["row_bytes"][0] == 514 * 4 and lay["plane_bytes"][0] == 514 * 514 * 4
    assert nb.value == 515 * 514 * 514 * 4
    from fvr._items import guard_offset
    assert guard_offset(lay[0]) == 513 * 514 * 514 * 4
    # the wrapped layout: one more index per axis
    assert L.fvr_layout_init(lay.ctypes.data, n.ctypes.data, 1 | (_lib.LAYOUT_WRAP << 8), ctypes.byref(nb),
                             ctypes.byref(ns)) == 0
    assert tuple(lay["sdim"][0]) == (516, 515, 515) and lay["flags"][0] == _lib.LAYOUT_WRAP
    assert lay["row_bytes"][0] == 515 * 4 and nb.value == 516 * 515 * 515 * 4
    # too large for 32-bit offsets -> error with a message, no abort
    big = np.array([1024, 1024, 1024], dtype=np.int32)
    assert L.fvr_layout_init(lay.ctypes.data, big.ctypes.data, 0, ctypes.byref(nb), ctypes.byref(ns)) == 1
    assert b"too large" in L.fvr_last_error()
    with pytest.raises(_lib.FvrNativeError):
        _lib.check(1, "probe")


def test_entry_points_reject_bad_arguments_without_a_gpu():
    from fvr import _lib
    L = _lib.lib()
    assert L.fvr_sncc_eval_v(None, 0, None, 0, 0, None, 0, None) == 1
    assert b"null" in L.fvr_last_error()
    assert L.fvr_slice_fwd(0, 0, 1, 1, 1, 1, None, None, 1, None, None, 1, 0.0, 0.0, 0.0, 0, None) == 1
    # a plan with a tile height the library has no kernel for, and a map-less slot count that
    # does not cover the tiles, are refused before any launch (fake non-null pointers)
    p = ctypes.c_void_p(16)
    assert L.fvr_sncc_eval_plan(p, 1, p, 1, 8, p, 0 | (9 << 4) | (96 << 8), None, 8, 0, None) == 1
    assert b"tile_h 96" in L.fvr_last_error()
    assert L.fvr_sncc_eval_plan(p, 1, p, 1, 9, p, 0 | (9 << 4) | (64 << 8), None, 8, 0, None) == 1
    assert b"slots" in L.fvr_last_error()
    assert L.fvr_sncc_eval_plan(p, 1, p, 1, 8, p, 0 | (9 << 4) | (64 << 8), None, 8, 9, None) == 1
    assert b"grid" in L.fvr_last_error()
    cfg = np.zeros(1, dtype=_lib.ADAM_DTYPE)
    assert L.fvr_adam_step(None, None, None, None, None, None, None, None, None, 1, cfg.ctypes.data, None) == 1
    cfg["beta1"] = 1.0
    assert L.fvr_adam_step(p, p, p, p, p, p, p, p, p, 1, cfg.ctypes.data, None) == 1
    assert b"hyper" in L.fvr_last_error()


def test_pou_packed_size_and_argument_checks():
    from fvr import _lib
    L = _lib.lib()
    # (P + 3) x (P + 2)^2 elements of 8 bytes for every dtype: a float2, or the 16-bit
    # (re, im) words of two mid-axis rows
    assert L.fvr_pou_packed_bytes(2, 55, 128) == 55 * 131 * 130 * 130 * 8
    assert L.fvr_pou_packed_bytes(1, 3, 32) == 3 * 35 * 34 * 34 * 8
    assert L.fvr_pou_packed_bytes(0, 1, 16) == 19 * 18 * 18 * 8
    assert L.fvr_pou_packed_bytes(3, 1, 16) == -1 and L.fvr_pou_packed_bytes(1, 0, 16) == -1
    assert L.fvr_pou_pack_blocks(0, 1, 1, 16, 0, None) == 1
    p = ctypes.c_void_p(16)
    assert L.fvr_pou_pack_blocks(16, 1, 600, 128, 16, None) == 1 and b"grid" in L.fvr_last_error()


def test_pou_gradient_scale_formula():
    """The fixed-point scale of the PoU image-gradient scatter (the function the gscale
    kernel evaluates): 2^(40 - e) for a maximum m 2^e, m in [0.5, 1); capped at 2^126 so
    that tiny and subnormal maxima keep a finite power of two (their fixed-point values stay
    below 2^40); 1 for an all-zero gradient; NaN for NaN or infinite maxima."""
    import math
    from fvr import _lib
    f = _lib.lib().fvr_pou_scale_of_max
    for mx in (1.0, 0.75, 3.0, 1e-3, 1e6, 2.0 ** -80):
        m, e = math.frexp(mx)
        assert f(mx) == 2.0 ** (40 - e), mx
        assert 2.0 ** 39 <= mx * f(mx) < 2.0 ** 40
    for mx in (2.0 ** -90, 1e-30, 2.0 ** -130, 1e-45):  # tiny and subnormal: the cap
        s = f(mx)
        assert s == 2.0 ** 126 and math.isfinite(s) and mx * s < 2.0 ** 40, mx
    assert f(0.0) == 1.0
    assert math.isnan(f(float("nan"))) and math.isnan(f(float("inf")))
    assert f(3.4e38) == 2.0 ** (40 - 128)


def test_kernel_variant_table():
    """Tile heights 32 / 128 exist for the batched adjoint forms; every flag set has 64."""
    from fvr import _lib
    L = _lib.lib()
    A, HP, WS, HERM = _lib.ITEM_ADJOINT, _lib.ITEM_HP, _lib.ITEM_WRITE_SLICE, _lib.ITEM_HERM
    for f in range(16):
        folded_ok = not (f & HERM and f & (HP | WS))
        assert L.fvr_sncc_has_tile_h(f, 64) == int(folded_ok)
        for th in (32, 128):
            want = folded_ok and (bool(f & HERM) or (bool(f & A) and not f & WS))
            assert L.fvr_sncc_has_tile_h(f, th) == int(want), (f, th)
        assert L.fvr_sncc_has_tile_h(f, 48) == 0
        for th in (32, 64, 128):
            assert L.fvr_sncc_has_depth(f, th, 2) == L.fvr_sncc_has_tile_h(f, th)
            assert L.fvr_sncc_has_depth(f, th, 3) == int(bool(f & HERM) and folded_ok)
            assert L.fvr_sncc_has_depth(f, th, 4) == 0
    assert L.fvr_sncc_has_tile_h(16, 64) == 0


def test_tile_maps_cover_every_tile_once_and_split_contiguously_over_xcds():
    from fvr.sncc import item_major_entries, slot_map, target_major_entries
    tiles = [(8, 4), (8, 4)]
    for n in (1, 5, 64):
        for ent in (target_major_entries(n, tiles, 6), item_major_entries(n, tiles, 6)):
            assert len(ent) == n * 64
            got = {tuple(e) for e in ent}
            assert got == {(6 + 2 * p + v, t) for p in range(n) for v in range(2) for t in range(32)}
            m = slot_map(ent)
            assert len(m) % 8 == 0 and len(m) - len(ent) < 8
            live = m[m[:, 0] >= 0]
            assert sorted(map(tuple, live)) == sorted(map(tuple, ent))
            per = len(m) // 8
            for x in range(8):  # XCD x's slots hold one contiguous run of the entry sequence
                run = m[x::8]
                run = run[run[:, 0] >= 0]
                np.testing.assert_array_equal(run, ent[x * per: x * per + len(run)])
    # target-major order: one XCD's run of a 64-pose batch is one (view, tile row) band
    m = slot_map(target_major_entries(64, tiles))
    for x in range(8):
        run = m[x::8]
        assert len({int(i) % 2 for i in run[:, 0]}) == 1 and len({int(t) // 8 for t in run[:, 1]}) == 1


def test_product_path_refuses_cpu():
    from fvr import _lib
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        _lib.default_device()
    with pytest.raises(RuntimeError):
        _lib.require_device(torch.zeros(3), "probe")
    from fvr import DifferentiableFourierVolumeRenderer
    with pytest.raises(RuntimeError):
        DifferentiableFourierVolumeRenderer(np.zeros((8, 8, 8)))


def test_item_builder_identity_geometry():
    """Host item geometry at the identity pose equals the integer slice grid (diff_fvr.py:128-139)."""
    from fvr import _items

    class FakePacked:
        shape = (16, 20, 24)
        index = {"zyx": 0, "zxy": 1}

        def layout_for(self, view, herm=False):
            return (1, (0, 2, 1)) if view == "lateral" else (0, (0, 1, 2))

        def desc(self, k):
            d = np.zeros(1, dtype=__import__("fvr")._lib.LAYOUT_DTYPE)[0]
            n = [16, 24, 20] if k == 1 else [16, 20, 24]
            d["n"] = n
            d["sdim"] = [n[0] + 3, n[1] + 2, n[2] + 2]
            d["plane_bytes"] = 1000
            return d

    it = _items.n_item(FakePacked(), "ap", 0, 0, 0, 0, 0, 0, 0, 0, 64)[0]
    assert (it["h"], it["w"]) == (16, 24) and it["tile_h"] == 64
    # c = (pz/2, py/2, px/2) + uu0*u + vv0*v with u = x, v = z, uu0 = -12, vv0 = -8
    # (every coordinate carries the +SNAP_FX lattice bias, and the box edge with it)
    np.testing.assert_array_equal(it["c"], (np.array([0, 10, 0]) << 32) + _items.SNAP_FX)
    np.testing.assert_array_equal(it["ustep"], np.array([0, 0, 1]) << 32)
    np.testing.assert_array_equal(it["vstep"], np.array([1, 0, 0]) << 32)
    lat = _items.n_item(FakePacked(), "lateral", 0, 0, 0, 0, 0, 0, 0, 0, 64)[0]
    # lateral uses the (z, x, y) layout: u = y is the fast axis
    np.testing.assert_array_equal(lat["ustep"], np.array([0, 0, 1]) << 32)
    np.testing.assert_array_equal(lat["c"], (np.array([0, 12, 0]) << 32) + _items.SNAP_FX)
    assert lat["box_hi"][2] == ((20 - 1) << 32) + _items.SNAP_FX
    # phase turns: a = 3 voxels over W = 24 -> 1/8 turn per column, uu0 * a / W = -1.5 -> 0.5 turn
    it = _items.n_item(FakePacked(), "ap", 0, 0, 0, 3.0, 0, 0, 0, 0, 64)[0]
    assert it["ph_col"] == 1 << 29 and it["ph_col0"] == 1 << 31


def test_jacobian_matches_finite_difference():
    from fvr import _items
    for view, perm in (("ap", (0, 1, 2)), ("lateral", (0, 2, 1))):
        y, p, r = 3.0, -7.0, 11.0
        du, dv = _items.n_jacobian(y, p, r, view, perm)
        h = 1e-6
        for j in range(3):
            a = [y, p, r]
            b = [y, p, r]
            a[j] += h
            b[j] -= h
            ua, va = _items.euler_to_basis_n(*a, view)
            ub, vb = _items.euler_to_basis_n(*b, view)
            np.testing.assert_allclose(du[:, j], ((ua - ub) / (2 * h))[list(perm)], atol=1e-9)
            np.testing.assert_allclose(dv[:, j], ((va - vb) / (2 * h))[list(perm)], atol=1e-9)


def test_phantom_is_seeded_and_in_range():
    from fvr.phantom import ellipsoid_phantom
    a = ellipsoid_phantom(24, seed=1234)
    b = ellipsoid_phantom(24, seed=1234)
    assert a.dtype == np.float32 and a.shape == (24, 24, 24)
    np.testing.assert_array_equal(a, b)
    assert a.max() > 0.0 and a.min() == 0.0
    assert not np.array_equal(a, ellipsoid_phantom(24, seed=1235))


def test_every_struct_field_offset_matches_the_header(tmp_path):
    """Compile a probe against include/fvr_hip.h and compare offsetof() of every field."""
    import shutil
    import subprocess
    from fvr import _lib
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, dt in (("fvr_sncc_item", _lib.ITEM_DTYPE), ("fvr_layout", _lib.LAYOUT_DTYPE)):
        for f in dt.names:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
        lines.append(f'printf("{cname}.__size__ %zu\\n", sizeof({cname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.split("\n") if l)
    for cname, dt in (("fvr_sncc_item", _lib.ITEM_DTYPE), ("fvr_layout", _lib.LAYOUT_DTYPE)):
        for f in dt.names:
            assert int(out[f"{cname}.{f}"]) == dt.fields[f][1], (cname, f)
        assert int(out[f"{cname}.__size__"]) == dt.itemsize


def test_t_target_and_pose_checks_reject_bad_inputs_before_any_launch():
    """The fused convention-T launch forms a target address per item from data_ptr() and
    stride(0): anything but (Bt, R, R) complex64 with contiguous rows, Bt in {1, B}, on the
    evaluator's device, is refused first (checked here against the CPU device)."""
    from fvr.torch_engine import check_t_poses, check_t_targets
    cpu = torch.device("cpu")
    good = torch.zeros((3, 8, 8), dtype=torch.complex64)
    assert check_t_targets(good, 3, cpu) == 8 and check_t_targets(good[:1], 3, cpu) == 8
    bad = [good.to(torch.complex128), good.real.contiguous(), good[:2], good[0], good[:, :, :4],
           good.transpose(1, 2), good[:, ::2, ::2], torch.zeros((3, 8, 16), dtype=torch.complex64)]
    for t in bad:
        with pytest.raises(ValueError):
            check_t_targets(t, 3, cpu)
    with pytest.raises(ValueError):
        check_t_targets(good, 3, torch.device("cuda", 0))
    P = torch.zeros((3, 6), dtype=torch.float64)
    assert check_t_poses(P, cpu) == 3
    for p in (P.float(), P[:, :5], P.t().contiguous(), torch.zeros((0, 6), dtype=torch.float64), P[::2]):
        with pytest.raises(ValueError):
            check_t_poses(p, cpu)
