"""CPU checks of the host-side set-up around the fused kernel (no GPU compute):
the vectorised item builder, the descriptor validator and the biplanar combine.
"""

import ctypes

import numpy as np
import pytest

from oracle import diff_fvr_np as O


class _Packed:
    """A PackedSpectrum stand-in with real layout descriptors (no device storage)."""

    def __init__(self, shape):
        from fvr import _lib
        L = _lib.lib()
        self.shape = shape
        Z, Y, X = shape
        descs = np.zeros(2, dtype=_lib.LAYOUT_DTYPE)
        for k, n in enumerate(([Z, Y, X], [Z, X, Y])):
            one = np.zeros(1, dtype=_lib.LAYOUT_DTYPE)
            nb, ns = ctypes.c_uint64(), ctypes.c_uint64()
            assert L.fvr_layout_init(one.ctypes.data, np.array(n, np.int32).ctypes.data, 0, ctypes.byref(nb),
                                     ctypes.byref(ns)) == 0
            descs[k] = one[0]
        self.descs = descs
        self.index = {"zyx": 0, "zxy": 1}

    def layout_for(self, view, herm=False):
        return (1, (0, 2, 1)) if view == "lateral" else (0, (0, 1, 2))

    def desc(self, k):
        return self.descs[k]


POSES = np.array([[3.3, 1.1, -2.2, 2.1, 1.5, -1.0], [-14.0, 13.0, 9.0, -18.0, 19.0, 7.0], [0.0] * 6,
                  [45.0, -30.0, 60.0, 0.25, -0.5, 3.0]])


@pytest.mark.parametrize("shape", [(48, 48, 48), (34, 40, 46)])
def test_vectorised_builder_equals_per_item_builder(shape):
    from fvr import _items
    pk = _Packed(shape)
    tm = np.concatenate([_items.n_template(pk, "ap", 1234, 1, 64), _items.n_template(pk, "lateral", 5678, 1, 64)])
    items, jac, n_tiles = _items.n_biplanar_items(tm, shape, POSES, (0, 2, 1), 64)
    assert len(items) == 2 * len(POSES)
    base = 0
    for k, p in enumerate(POSES):
        for v, view in enumerate(("ap", "lateral")):
            a = p[3] if view == "ap" else p[4]
            ref = _items.n_item(pk, view, p[0], p[1], p[2], a, p[5], 1234 if v == 0 else 5678, 1, base, 64)[0]
            got = items[2 * k + v]
            for f in ref.dtype.names:
                np.testing.assert_array_equal(got[f], ref[f], err_msg=f"{view} pose {k} field {f}")
            du, dv = _items.n_jacobian(p[0], p[1], p[2], view, pk.layout_for(view)[1])
            np.testing.assert_allclose(jac[2 * k + v, 0], du, atol=1e-15)
            np.testing.assert_allclose(jac[2 * k + v, 1], dv, atol=1e-15)
            base += int(np.prod(_items.tiles_for(int(ref["h"]), int(ref["w"]), 64)))
    assert n_tiles == base
    assert _items.validate_items(items, 2, pk.descs, 64) == n_tiles


def test_validator_rejects_malformed_descriptors():
    from fvr import _items
    pk = _Packed((32, 32, 32))
    tm = np.concatenate([_items.n_template(pk, "ap", 1, 1, 64), _items.n_template(pk, "lateral", 2, 1, 64)])
    items, _, _ = _items.n_biplanar_items(tm, pk.shape, POSES[:2], (0, 2, 1), 64)
    for field, value in (("layout", 2), ("layout", -1), ("lo_int", 2), ("target", 0), ("h", 0),
                         ("guard_off", 0), ("tile_base", 7), ("tiles_x", 5)):
        bad = items.copy()
        bad[1][field] = value
        with pytest.raises(ValueError):
            _items.validate_items(bad, 2, pk.descs, 64)
    bad = items.copy()
    bad[0]["box_hi"][0] = np.uint64(40) << np.uint64(32)
    with pytest.raises(ValueError):
        _items.validate_items(bad, 2, pk.descs, 64)
    bad = items.copy()
    bad[0]["flags"] = 4  # slice output without a buffer
    with pytest.raises(ValueError):
        _items.validate_items(bad, 2, pk.descs, 64)


def _moments_from_oracle(f3, pose, view, T, perm):
    """The kernel's 16 moments computed directly in float64 from oracle I, dI (layout axes)."""
    yaw, pitch, roll, tx, ty, tz = pose
    a = tx if view == "ap" else ty
    h, w = O.slice_dims(f3.shape, view)
    u, v = O.euler_to_basis(yaw, pitch, roll, view)
    coords, uu, vv = O.slice_coords(f3.shape, u, v, view)
    I, dI = O.trilinear_value_and_grad(f3, coords)
    uu, vv = uu.ravel(), vv.ravel()
    ph = np.exp(-2j * np.pi * (uu * a / w + vv * tz / h))
    W = ph * np.conj(T.ravel())
    IW = I * W
    g = np.real(dI * W[None])[list(perm)]
    hh = np.real(dI * np.conj(I)[None])[list(perm)]
    m = np.zeros(16)
    m[0], m[1] = IW.real.sum(), np.sum(np.abs(I) ** 2)
    m[2], m[3] = np.sum(uu * IW.imag), np.sum(vv * IW.imag)
    m[4:7], m[7:10] = (g * uu).sum(1), (g * vv).sum(1)
    m[10:13], m[13:16] = (hh * uu).sum(1), (hh * vv).sum(1)
    return m


def test_combine_reproduces_oracle_gradient():
    from fvr import _items
    from fvr.sncc import combine_biplanar
    vol = np.zeros((16, 16, 16))
    vol[4:11, 5:12, 3:9] = 1.0
    vol[9:13, 2:6, 8:14] = 0.5
    f3 = O.make_spectrum(vol)
    tp = [1.0, -2.0, 1.5, 0.7, -1.2, 0.4]
    T_ap = O.extract_frequency_slice(f3, *tp[:3], tx=tp[3], ty=tp[5], view="ap")
    T_la = O.extract_frequency_slice(f3, *tp[:3], tx=tp[4], ty=tp[5], view="lateral")
    poses = np.array([[3.3, 1.1, -2.2, 2.1, 1.5, -1.0], [-6.0, 4.0, 2.0, -3.0, 0.5, 2.5]])
    pk = _Packed(f3.shape)
    mom = np.stack([_moments_from_oracle(f3, p, view, T, perm)
                    for p in poses for view, T, perm in (("ap", T_ap, (0, 1, 2)), ("lateral", T_la, (0, 2, 1)))])
    tm = np.concatenate([_items.n_template(pk, "ap", 1, 1, 64), _items.n_template(pk, "lateral", 2, 1, 64)])
    _, jac, _ = _items.n_biplanar_items(tm, f3.shape, poses, (0, 2, 1), 64)
    t2 = [np.sum(np.abs(T_ap) ** 2), np.sum(np.abs(T_la) ** 2)]
    hw = {"ap": O.slice_dims(f3.shape, "ap"), "lateral": O.slice_dims(f3.shape, "lateral")}
    loss, grad = combine_biplanar(mom, jac, t2, hw)
    for k, p in enumerate(poses):
        Lr, gr = O.biplanar_loss_and_grad_analytic(f3, p, T_ap, T_la)
        assert abs(loss[k] - Lr) < 1e-12
        np.testing.assert_allclose(grad[k], gr, rtol=1e-9, atol=1e-14)
    # zero target -> NCC 0, loss 1, gradient 0 (diff_fvr.py:193-194)
    loss0, grad0 = combine_biplanar(mom, jac, [0.0, 0.0], hw)
    assert np.all(loss0 == 1.0) and np.all(grad0 == 0.0)


def test_footprint_counter_equals_direct_count():
    from fvr.footprint import slice_footprint
    shape, pose = (40, 40, 40), (7.0, -4.0, 3.0)
    U, S = slice_footprint(shape, pose, "lateral")
    assert S == shape[0] * shape[1]
    u, v = O.euler_to_basis(*pose, "lateral")
    c, _, _ = O.slice_coords(shape, u, v, "lateral")
    n = np.array(shape)[:, None]
    c = c[:, np.all((c >= 0) & (c <= n - 1), 0)]
    i0 = np.floor(c).astype(int)
    t = c - i0
    keys = set()
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (t[0] if dz else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dx else 1 - t[2])
                idx = i0[:, w > 0] + np.array([dz, dy, dx])[:, None]
                keys |= set(map(tuple, idx.T))
    assert U == len(keys)
    assert 1.9 < U / (shape[0] * shape[1]) < 2.6


def test_item_byte_image_keeps_the_c_layout():
    """Regression: NumPy concatenation of structured items must not shift fields.

    A packed item dtype (alignment hole dropped) once moved ``target`` from byte
    184 to 180, so the kernel read its target address from the wrong bytes.
    """
    from fvr import _items, _lib
    a = np.zeros(1, dtype=_lib.ITEM_DTYPE)
    c = np.concatenate([a, a])
    assert c.dtype.itemsize == 200 and c.dtype.fields["target"][1] == 184
    pk = _Packed((32, 32, 32))
    tm = np.concatenate([_items.n_template(pk, "ap", 0x1111222233334444, 1, 64),
                         _items.n_template(pk, "lateral", 0x5555666677778888, 1, 64)])
    items, _, _ = _items.n_biplanar_items(tm, pk.shape, POSES[:3], (0, 2, 1), 64)
    raw = _lib.struct_bytes(items, _lib.ITEM_DTYPE)
    assert raw.size == 200 * len(items)
    for k in range(len(items)):
        tgt = int.from_bytes(raw[200 * k + 184: 200 * k + 192].tobytes(), "little")
        assert tgt == (0x1111222233334444 if k % 2 == 0 else 0x5555666677778888)
        assert int.from_bytes(raw[200 * k + 156: 200 * k + 160].tobytes(), "little") == (0 if k % 2 == 0 else 1)
    # a packed (196-byte) copy is converted back by field name
    packed = np.zeros(2, dtype=np.dtype([(n, _lib.ITEM_DTYPE.fields[n][0]) for n in _lib.ITEM_DTYPE.names]))
    packed["target"] = [7, 9]
    assert _lib.as_struct(packed, _lib.ITEM_DTYPE)["target"].tolist() == [7, 9]
    assert _lib.struct_bytes(packed, _lib.ITEM_DTYPE).size == 400


def test_pou_slice_footprint_closed_form_and_brute_force():
    """fvr.footprint.pou_slice_footprint: an axis-aligned slab-image slice samples lattice
    points only (one corner with weight 1 each: U = P^2); at a tilted pose the count equals
    a sample-by-sample loop over the trilinear corners inside [0, P - 1]^3; and
    pou_step_bytes adds the kernels' bytes as its docstring states."""
    from fvr.footprint import pou_slice_footprint, pou_step_bytes
    P = 8
    assert pou_slice_footprint(P, [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]) == P * P
    Q = np.linalg.qr(np.random.default_rng(3).standard_normal((3, 3)))[0]
    u, v = Q[:, 0], Q[:, 1]
    keys = set()
    for i in range(P):
        for j in range(P):
            p = P // 2 + (j - P // 2) * u + (i - P // 2) * v
            i0 = np.floor(p).astype(int)
            t = p - i0
            for d in np.ndindex(2, 2, 2):
                w = np.prod([t[a] if d[a] else 1.0 - t[a] for a in range(3)])
                c = i0 + np.array(d)
                if w > 0 and np.all((c >= 0) & (c <= P - 1)):
                    keys.add(tuple(c))
    U = pou_slice_footprint(P, u, v)
    assert U == len(keys) and P * P < U < 8 * P * P
    B, K, tex = 3, 100, 1234
    fb = pou_step_bytes(B, P, K, np.stack([u, u]), np.stack([v, v]), 4, texels_read=tex)
    g, img, rays, out = 2 * U * B * 4, tex * 4, 3 * K * 4, 2 * K * 4
    assert (fb["slice_ifft"], fb["warp_fwd"], fb["render_bwd"]) == (g + img, img + out + rays, g + img + out + rays)
    assert fb["step"] == 2 * g + 3 * img + 2 * out + 2 * rays and fb["slab_texels_read"] == tex
    assert pou_step_bytes(B, P, K, np.stack([u]), np.stack([v]), 4)["slab_texels_read"] == B * P * P


def test_hermitian_footprint_is_about_half():
    """The fold schedule touches about half of the full slice's distinct entries and
    evaluates about half of its samples."""
    from fvr.footprint import slice_footprint
    shape, pose = (40, 40, 40), (7.0, -4.0, 3.0)
    U, S = slice_footprint(shape, pose, "ap")
    Uh, Sh = slice_footprint(shape, pose, "ap", herm=True)
    h, w = 40, 40
    assert Sh == (h // 2 - 1) * w + (1 + w // 2) + (min(h, w) // 2 - 1) + w
    assert 0.45 < Uh / U < 0.6 and 0.45 < Sh / S < 0.6


def test_product_mtre_reference_closed_forms_and_oracle():
    """fvr.registration_3d2d.compute_mtre_6dof (registration_3d2d.py:112-175): the
    reference's known answers (tests/test_physical_units.py:52-59: 6.0 for a 3-voxel shift
    at 2 mm, 3.0 for 3 mm, 16.0 for a 90 degree yaw of the 8 landmarks at r = 4 voxels,
    2 mm), a 180 degree roll (every landmark moves by 2 r in its two rotated coordinates:
    2 * sqrt(2) * 8 mm), and the oracle's restatement on random poses and shapes."""
    from fvr.registration_3d2d import compute_mtre_6dof
    o = (0.0,) * 6
    assert abs(compute_mtre_6dof(o, (0, 0, 0, 3, 0, 0), (16,) * 3, spacing=(2.0,) * 3) - 6.0) < 1e-12
    assert abs(compute_mtre_6dof(o, (0, 0, 0, 3, 0, 0), (16,) * 3, spacing=(2.0,) * 3,
                                 translation_units="mm") - 3.0) < 1e-12
    assert abs(compute_mtre_6dof(o, (90, 0, 0, 0, 0, 0), (16,) * 3, spacing=(2.0,) * 3) - 16.0) < 1e-12
    assert abs(compute_mtre_6dof(o, (0, 0, 180, 0, 0, 0), (16,) * 3, spacing=(2.0,) * 3)
               - 2.0 * np.sqrt(2.0) * 8.0) < 1e-12
    rng = np.random.default_rng(5)
    for _ in range(20):
        a, b = rng.uniform(-20, 20, 6), rng.uniform(-20, 20, 6)
        shape = tuple(int(x) for x in rng.integers(8, 300, 3))
        sp = tuple(rng.uniform(0.3, 3.0, 3))
        for units in ("voxels", "mm"):
            got = compute_mtre_6dof(a, b, shape, spacing=sp, translation_units=units)
            ref = O.compute_mtre_6dof(a, b, shape, spacing=sp, translation_units=units)
            assert abs(got - ref) <= 1e-12 * max(1.0, ref)
    for sp in ((0, 1, 1), (1.0, float("nan"), 1.0), (1.0, 1.0)):
        with pytest.raises(ValueError):
            compute_mtre_6dof(o, o, (16,) * 3, spacing=sp)
    with pytest.raises(ValueError):
        compute_mtre_6dof(o, o, (16,) * 3, translation_units="cm")


def test_pn_evaluate_workspace_size_and_argument_checks():
    """fvr_pn_evaluate's workspace query (items | Jacobians | tile partials, 256-byte
    aligned) and its argument checks, which reject before any launch (no GPU here)."""
    from fvr import _items, _lib
    L = _lib.lib()
    pk = _Packed((48, 48, 48))
    flags = _lib.ITEM_ADJOINT
    tm = _lib.as_struct(np.concatenate([_items.n_template(pk, "ap", 1234, flags, 64),
                                        _items.n_template(pk, "lateral", 5678, flags, 64)]), _lib.ITEM_DTYPE)
    tiles = sum(int(np.prod(_items.tiles_for(int(t["h"]), int(t["w"]), 64))) for t in tm)

    def a256(b):
        return (b + 255) // 256 * 256

    for n in (1, 7, 64):
        exp = a256(2 * n * _lib.ITEM_DTYPE.itemsize) + a256(2 * n * 18 * 8) + a256(n * tiles * 16 * 4)
        assert L.fvr_pn_evaluate_workspace(tm.ctypes.data, n) == exp
    assert L.fvr_pn_evaluate_workspace(tm.ctypes.data, 0) == -1
    assert L.fvr_pn_evaluate_workspace(None, 4) == -1
    fake = ctypes.c_void_p(1 << 20)  # never dereferenced: every call below fails its checks first
    var = 1 << 4 | 64 << 8  # f32 storage, adjoint, tile_h 64
    need = L.fvr_pn_evaluate_workspace(tm.ctypes.data, 4)

    def call(tmpl=tm, variant=var, ws=fake, nbytes=need, n=4, dims=(48.0, 48.0, 48.0)):
        return L.fvr_pn_evaluate(fake, 2, tmpl.ctypes.data, *dims, 1, variant, fake, n, 1.0, 1.0, ws, nbytes,
                                 None, fake, fake, None)

    assert call(nbytes=need - 1) == 1 and b"workspace" in L.fvr_last_error()
    assert call(ws=ctypes.c_void_p((1 << 20) + 64)) == 1 and b"aligned" in L.fvr_last_error()
    assert call(variant=1 << 4 | 32 << 8) == 1 and b"tile_h" in L.fvr_last_error()
    assert call(variant=0 | 64 << 8) == 1 and b"flags" in L.fvr_last_error()
    assert call(n=0) == 1
    assert call(dims=(48.0, 48.0, 46.0)) == 1 and b"padded dims" in L.fvr_last_error()


def test_autotune_and_biplanar_order_argument_checks():
    """fvr_autotune_pn_evaluate rejects bad arguments before any HIP call (as
    fvr_pn_evaluate does, plus its own: repetitions, an unaligned workspace, missing
    outputs); fvr_autotune_clear needs no device; the fused kernel's biplanar tile order
    (variant bit 21) refuses a tile map and an odd item count, and bits above 21 are
    refused (no GPU here)."""
    from fvr import _items, _lib
    L = _lib.lib()
    pk = _Packed((48, 48, 48))
    flags = _lib.ITEM_ADJOINT
    tm = _lib.as_struct(np.concatenate([_items.n_template(pk, "ap", 1234, flags, 64),
                                        _items.n_template(pk, "lateral", 5678, flags, 64)]), _lib.ITEM_DTYPE)
    fake = ctypes.c_void_p(1 << 20)  # never dereferenced: every call below fails its checks first
    var = 1 << 4 | 64 << 8
    need = L.fvr_pn_evaluate_workspace(tm.ctypes.data, 4)
    best, ms = ctypes.c_uint32(0), ctypes.c_float(0.0)

    def tune(variant=var, ws=fake, reps=3, out=True, n=4):
        return L.fvr_autotune_pn_evaluate(fake, 2, tm.ctypes.data, 48.0, 48.0, 48.0, 1, variant, fake, n, 1.0, 1.0,
                                          ws, need, fake, fake, reps, 1, ctypes.byref(best) if out else None,
                                          ctypes.byref(ms), None)

    assert tune(reps=0) == 1 and b"bad arguments" in L.fvr_last_error()
    assert tune(out=False) == 1
    assert tune(ws=ctypes.c_void_p((1 << 20) + 8)) == 1 and b"aligned" in L.fvr_last_error()
    assert tune(variant=1 << 4 | 128 << 8) == 1 and b"tile_h" in L.fvr_last_error()
    assert tune(n=0) == 1
    assert best.value == 0
    assert L.fvr_autotune_clear() == 0

    def plan(variant, n_items=2, tile_map=None):
        return L.fvr_sncc_eval_plan(fake, 2, fake, n_items, 8, fake, variant, tile_map, 8, 0, None)

    assert plan(var | 1 << 21, tile_map=fake) == 1 and b"biplanar" in L.fvr_last_error()
    assert plan(var | 1 << 21, n_items=3) == 1 and b"biplanar" in L.fvr_last_error()
    assert plan(var | 1 << 22) == 1 and b"bad variant" in L.fvr_last_error()


def test_pn_evaluate_noalloc_rejects_batches_beyond_its_scratch_before_any_device_call():
    """fvr_pn_evaluate_noalloc's size limit (the static scratch array of
    FVR_PN_SCRATCH_BYTES): a batch whose fvr_pn_evaluate_workspace exceeds it is refused
    with a message naming the limit, before any HIP call (no GPU here)."""
    from fvr import _items, _lib
    L = _lib.lib()
    pk = _Packed((48, 48, 48))
    flags = _lib.ITEM_ADJOINT
    tm = _lib.as_struct(np.concatenate([_items.n_template(pk, "ap", 1234, flags, 64),
                                        _items.n_template(pk, "lateral", 5678, flags, 64)]), _lib.ITEM_DTYPE)
    n = 64
    while L.fvr_pn_evaluate_workspace(tm.ctypes.data, n) <= (16 << 20):
        n *= 2
    fake = ctypes.c_void_p(1 << 20)
    rc = L.fvr_pn_evaluate_noalloc(fake, 2, tm.ctypes.data, 48.0, 48.0, 48.0, 1, 1 << 4 | 64 << 8, fake, n, 1.0, 1.0,
                                   None, fake, fake, None)
    assert rc == 1 and b"scratch" in L.fvr_last_error() and str(16 << 20).encode() in L.fvr_last_error()


def test_pinhole_lattice_ray_cache_is_exactly_separable_for_a_zero_skew_pinhole():
    """make_pinhole_lattice forms the ray cache in float64: for a zero-skew pinhole every row
    of r_x is row 0's and every column of r_y column 0's, bit for bit (the condition of the
    gather form of the PoU warp adjoint), and the values are the oracle's float64 cache
    rounded once; a skewed pinhole keeps its full arrays."""
    import torch
    from fvr import PerspectiveSlabFVR
    from oracle import cone_beam_np as C
    H, W, f, sdd = 20, 22, 300.0, 600.0
    k_inv = torch.tensor([[1.0 / f, 0.0, -W / (2 * f)], [0.0, 1.0 / f, -H / (2 * f)], [0.0, 0.0, 1.0]])
    lat = PerspectiveSlabFVR.make_pinhole_lattice(k_inv, sdd, H, W)
    rx, ry, ic = lat[5:8]
    assert rx.dtype == ry.dtype == ic.dtype == torch.float32 and tuple(rx.shape) == (H, W)
    assert torch.equal(rx, rx[:1].expand(H, W)) and torch.equal(ry, ry[:, :1].expand(H, W))
    assert torch.equal(lat[8], rx.reshape(-1)) and torch.equal(lat[9], ry.reshape(-1))
    ref = C.make_pinhole_lattice(k_inv.double().numpy(), sdd, H, W)
    for got, key in ((rx, "rx"), (ry, "ry"), (ic, "ic")):
        assert np.max(np.abs(got.double().numpy() - ref[key])) <= 1e-7 * np.abs(ref[key]).max()
    k_skew = k_inv.clone()
    k_skew[0, 1] = 2e-4
    rs = PerspectiveSlabFVR.make_pinhole_lattice(k_skew, sdd, H, W)[5]
    assert not torch.equal(rs, rs[:1].expand(H, W))
