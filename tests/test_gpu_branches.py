"""HIP parity on the branches the headline payloads do not reach (needs an MI355X).

* point location on hand-built trellises: node-subscript neighbour steps
  (src/trellis_poly.hpp:382-434) and the PolyNode fallback
  (src/trellis_node.hpp:275-308), every aux field bit for bit against the oracle;
* real-lattice vector, axial vector and matrix rotations (rip_real, rip_axial)
  through the general sums, FP64 within 1e-12 of the oracle and FP32 within
  1e-5, and against the closed forms of tests/test_oracle_rotations.py;
* Gamma rotation of Cartesian (angstrom) eigenvectors on the NaCl fixture;
* a non-identity PermutationTable passed through the C ABI's perm descriptor.
"""
import sys

import numpy as np
import pytest

from brille_amd import capi
from conftest import (CUBE, J_Z, POLY, cartesian, direct_components, direct_operator, field_grid, fill_field,
                      hand_trellis, nacl_grid, phase_free_vectors)
from oracle import oracle as O
from test_oracle_locate import A0, A1, SMALL, _poly_tables

pytestmark = pytest.mark.gpu

LOCATE_AUX = ("node", "tet", "n_iw", "iw_index", "iw_weight", "status")
KERNELS = [capi.KERNEL_PLAIN, capi.KERNEL_TILE, capi.KERNEL_BINNED]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hand_cases():
    cases = [
        (hand_trellis({(1, 0, 0): CUBE}), [[2.0, 0.5, 0.5], [1.5, 0.5, 0.5]]),
        (hand_trellis({(0, 3, 0): CUBE}), [[0.5, 3.0 - 1e-13, 0.5], [0.5, 3.25, 0.5]]),
        (hand_trellis({(1, 1, 2): CUBE}), [[2.0, 2.0, 2.5]]),
        (hand_trellis({(1, 1, 1): CUBE}), [[2.0, 2.0, 2.0], [1.0, 1.0, 1.0]]),
        (hand_trellis({(3, 1, 1): CUBE}), [[4.0, 1.5, 1.5]]),
        (hand_trellis({(2, 1, 1): CUBE, (1, 1, 1): CUBE}), [[2.0, 1.5, 1.5]]),
        (hand_trellis({(0, 2, 2): CUBE, (1, 1, 2): CUBE}), [[2.0, 2.0, 2.5]]),
        (hand_trellis({(0, 0, 0): CUBE}), [[-0.25, 0.5, 0.5]]),
        (hand_trellis({(3, 3, 3): CUBE}), [[2.0, 0.5, 0.5], [2.5, 2.5, 2.5], [3.5, 3.5, 3.5]]),
        (_poly_tables([A0, A1]), [[1.5, 1.0, 1.0], [1.75, 1.625, 1.625], [1.5, 1.625, 1.625]]),
        (_poly_tables([A1, A0]), [[1.5, 1.0, 1.0], [1.5, 1.625, 1.625]]),
        (_poly_tables([A0, A1, SMALL]), [[1.75, 1.625, 1.625], [1.9, 1.9, 1.9]]),
    ]
    return cases


@pytest.mark.parametrize("kernel", KERNELS)
def test_hand_built_trellises_locate_like_the_oracle(kernel):
    for ci, (t, x) in enumerate(_hand_cases()):
        say(f"hand-built trellis case {ci} kernel={kernel}")
        x = np.asarray(x, float)
        dev = capi.Grid(t)
        orc, ov, oc, oaux, _ = O.ir_interpolate_at(t, x, no_move=True)
        rc, gv, gc, gaux = dev.run_device(x, precision=capi.FP64, kernel=kernel, no_move=True, check=False,
                                          aux=LOCATE_AUX)
        assert rc == 0
        ok = oaux["status"] == 0
        for k in LOCATE_AUX:
            a, b = gaux[k], oaux[k]
            if k in ("node", "tet", "n_iw", "iw_index", "iw_weight"):
                a, b = a[ok], b[ok]
            np.testing.assert_array_equal(a, b, err_msg=f"case {ci} {k}")
        np.testing.assert_allclose(gv[ok], ov[ok], rtol=0, atol=1e-14, err_msg=f"case {ci}")
        assert not gv[~ok].any()
        # with the check on, the first failing point is the error
        if (~ok).any():
            rc, *_ = dev.run_device(x, precision=capi.FP64, kernel=kernel, no_move=True, check=True, aux=())
            assert capi.STATUS_NAMES[rc] == "NULL_NODE"


VECTOR, PSEUDOVECTOR, REAL_LATTICE = 0, 1, 3
HEX = ("194", (3.2, 3.2, 5.2), (90, 90, 120))
HEX_POLAR = ("186", (3.2, 3.2, 5.2), (90, 90, 120))


def _rotation_payloads():
    """(name, case, values builder, elements, complex vectors builder, vector elements)."""
    def vec(lat):
        return lambda q: direct_components(lat, cartesian(lat, q))

    def axial(lat):
        return lambda q: direct_components(lat, cartesian(lat, q) @ J_Z.T)

    def op(lat):
        z = np.array([0.0, 0.0, 1.0])
        return lambda q: direct_operator(lat, np.einsum("ni,j->nij", cartesian(lat, q), z)
                                         + np.einsum("i,nj->nij", z, cartesian(lat, q)))

    def rows(f, ph):
        return lambda q: (np.hstack([np.full((q.shape[0], 1), 10.0), f(q)]) * ph)[:, None, :]

    return [
        ("real vectors", HEX, vec, [0, 3, 0, VECTOR, REAL_LATTICE], lambda f: rows(f, 1 + 0.5j),
         [1, 3, 0, VECTOR, REAL_LATTICE]),
        ("axial vectors", HEX_POLAR, axial, [0, 3, 0, PSEUDOVECTOR, REAL_LATTICE], lambda f: rows(f, 2 - 1j),
         [1, 3, 0, PSEUDOVECTOR, REAL_LATTICE]),
        ("real matrices", HEX_POLAR, op, [0, 0, 9, VECTOR, REAL_LATTICE], None, None),
        ("axial matrices", HEX_POLAR, op, [0, 0, 9, PSEUDOVECTOR, REAL_LATTICE], None, None),
    ]


@pytest.mark.parametrize("which", range(4))
def test_real_lattice_rotations_match_the_oracle(which):
    name, case, field, el, cvec, cel = _rotation_payloads()[which]
    say(f"rotation payload {name}")
    lat, bz, g = field_grid(*case)
    f = field(lat)
    rlu = np.asarray(g.rlu)
    if cvec is None:
        t = fill_field(g, f(rlu), el)
    else:
        t = fill_field(g, f(rlu), el, cvec(f)(rlu), cel)
    Q = np.random.default_rng(10 + which).uniform(-3, 3, size=(20000, 3))
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    assert rc == 0
    dev = capi.Grid(t)
    for kernel in KERNELS:
        gv, gc, gaux = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=kernel,
                                             aux=dict(tau=kernel != capi.KERNEL_BINNED))
        for k in gaux:
            np.testing.assert_array_equal(gaux[k], oaux[k], err_msg=f"{name} kernel {kernel} {k}")
        sv, sc = np.abs(ov).max(), max(np.abs(oc).max(), 1e-300)
        assert np.max(np.abs(gv - ov)) <= 1e-12 * sv, (name, kernel)
        assert np.max(np.abs(gc - oc)) <= 1e-12 * sc, (name, kernel)
        fv, fc = dev.ir_interpolate_at(Q, precision=capi.FP32, kernel=kernel)
        assert np.max(np.abs(fv - ov)) <= 1e-5 * sv, (name, kernel)
        assert np.max(np.abs(fc - oc)) <= 1e-5 * sc, (name, kernel)


def test_cartesian_gamma_rotation_on_the_nacl_fixture(nacl):
    """wrap/tests/test_5_gamma.py:180-221 on the GPU: Cartesian (angstrom) eigenvectors rotate by the
    Cartesian matrices of the GammaTable."""
    say("nacl cartesian gamma on gpu")
    d = nacl
    lat, bz, g, perm, q_nu, _ = nacl_grid(d)
    cart = np.einsum("ba,ijkb->ijka", d["basis_vectors"], d["grid_vectors"])
    g.fill(d["grid_values"][perm], d["grid_values_elements"].astype(int), d["grid_values_weights"],
           cart[perm], np.array([0, 24, 0, 2, 1, 0, 0]), d["grid_vectors_weights"], False)
    t = g.tables()
    Q = np.vstack([d["q_nu"], np.random.default_rng(3).uniform(-2, 2, size=(4000, 3))])
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    assert rc == 0
    dev = capi.Grid(t)
    n = len(d["q_nu"])
    for kernel in KERNELS:
        gv, gc = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=kernel)
        assert np.max(np.abs(gc - oc)) <= 1e-12 * np.abs(oc).max(), kernel
        assert np.max(np.abs(gv - ov)) <= 1e-12 * np.abs(ov).max(), kernel
        assert np.allclose(phase_free_vectors(gc[:n], d["euphonic_vectors"]), d["euphonic_vectors"])
        # FP32 on the fixture's own q (grid points): between grid points the NaCl modes are
        # degenerate, their Hermitian overlaps near zero, and the antiphase of such an overlap
        # is not determined to FP32 precision
        fv, fc = dev.ir_interpolate_at(d["q_nu"], precision=capi.FP32, kernel=kernel)
        assert np.allclose(phase_free_vectors(fc.astype(np.complex128), d["euphonic_vectors"]),
                           d["euphonic_vectors"], atol=1e-4), kernel


def test_permutation_table_through_the_c_abi():
    """PermutationTable::safe_get (src/permutation_table.hpp:193-198) with 2000 stored (pivot, vertex)
    permutations: the general sums follow them exactly like the oracle."""
    say("permutation table")
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    V, B = t["vertices"].shape[0], int(t["branches"])
    r = np.random.default_rng(21)
    Q = S.q_points(1, 20000, stream=21)
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    # permutations for pairs the points actually use (pivot = first kept vertex)
    piv = oaux["iw_index"][:, 0].astype(np.uint64)
    other = oaux["iw_index"][:, 1].astype(np.uint64)
    keys = np.unique(piv * V + other)[:2000]
    perms = np.array([r.permutation(B) for _ in range(17)], np.uint32)
    index = r.integers(0, len(perms), size=len(keys)).astype(np.uint32)
    table = (keys, index, perms)
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=8, perm=table)
    assert rc == 0
    rc0, pv, pc, _, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    assert np.abs(pv - ov).max() > 1e-3  # the table changes results
    dev = capi.Grid(t, perm=table)
    for kernel in KERNELS:
        gv, gc = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=kernel)
        assert np.max(np.abs(gv - ov)) <= 1e-12 * np.abs(ov).max(), kernel
        assert np.max(np.abs(gc - oc)) <= 1e-12 * np.abs(oc).max(), kernel
        fv, fc = dev.ir_interpolate_at(Q, precision=capi.FP32, kernel=kernel)
        assert np.max(np.abs(fv - ov)) <= 1e-5 * np.abs(ov).max(), kernel


def test_fill_rejects_out_of_range_gamma_tables():
    """brille_grid_fill checks every GammaTable entry the kernels index with before uploading:
    an atom mapping >= n_atoms or a vector index >= n_vectors is an invalid argument."""
    say("gamma table validation")
    import ctypes as C
    from brille_amd import synthetic as S
    t = S.grid(1).tables()
    dev = capi.Grid(t)
    lib = capi.lib()
    for key, bad in (("gamma_F0", lambda a: a.max() + 1), ("gamma_vector_index", lambda a: 10 ** 6)):
        tt = dict(t)
        arr = np.array(t[key], np.uint32).copy()
        arr.flat[5] = bad(arr)
        tt[key] = arr
        h = capi.DataHolder(tt)
        rc = lib.brille_grid_fill(dev.h, C.byref(h.vdesc), C.byref(h.cdesc), C.byref(h.gdesc), None)
        assert capi.STATUS_NAMES[rc] == "INVALID_ARGUMENT", key
    # the valid table still fills and runs
    dev.fill(t)
    v, c = dev.ir_interpolate_at(S.q_points(1, 100, stream=7))
    assert np.isfinite(v).all()
