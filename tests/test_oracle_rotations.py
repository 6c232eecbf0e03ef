"""The oracle's non-Gamma rotations pinned by closed-form fields.

Interpolator::rotate_in_place (src/interpolator.hpp:387-428) sends real-lattice
data to rip_real or rip_axial (src/interpolator_real.tpp:19-64,
src/interpolator_axial.tpp:19-64).  A field that is linear in q is reproduced
exactly by the trellis' linear interpolation, so after ir_interpolate_at the
result is a closed form of the field, the point's operation indices and its
folded q:

* rip_real, vectors:    W[Ridx] f(ir_q)
* rip_axial, vectors:   det(W[Ridx]) W[invRidx] f(ir_q)
* rip_real, matrices:   W[Ridx] M(ir_q) W[invRidx]
* rip_axial, matrices:  W[invRidx] M(ir_q) W[Ridx]

with W the real-space point-group matrices (direct basis) in PointSymmetry
order.  Where the reference's choice is the symmetry-consistent one the result
is also the field itself at Q - tau (equivariance), a check that does not use
the formula at all: for axial vectors and axial matrices always, for vectors
and matrices in real-lattice units only in point groups whose operations are
their own inverses (rip_real applies W[Ridx] where an equivariant field needs
W[invRidx]; for 3-, 4- and 6-fold axes the two differ, and the oracle follows
the reference).  Ridx/invRidx themselves are pinned by the wedge brute force and
the NaCl fixture (test_oracle.py).
"""
import numpy as np
import pytest

from conftest import J_Z, cartesian, direct_components, direct_operator, field_grid, fill_field
from oracle import oracle as O

VECTOR, PSEUDOVECTOR = 0, 1
REAL_LATTICE = 3

HEX = ("194", (3.2, 3.2, 5.2), (90, 90, 120))       # 6/mmm: 6-fold and 3-fold axes
HEX_POLAR = ("186", (3.2, 3.2, 5.2), (90, 90, 120))  # 6mm: fixes z, has mirrors (det -1)
ORTHO = ("62", (5.0, 6.0, 7.5), (90, 90, 90))        # mmm: every operation its own inverse
ORTHO_POLAR = ("25", (4.1, 5.3, 6.2), (90, 90, 90))  # mm2: fixes z, every operation its own inverse


def _run(t, Q):
    rc, v, c, aux, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    assert rc == 0
    return v[:, 0, :], c[:, 0, :], aux


def _w(t):
    return t["rotations"].reshape(-1, 3, 3).astype(float)


def _vector_field(lat):
    return lambda q: direct_components(lat, cartesian(lat, q))


def _axial_field(lat):
    return lambda q: direct_components(lat, cartesian(lat, q) @ J_Z.T)


def _operator_field(lat):
    """M(q) = x z^T + z x^T in the direct basis: an operator field for groups that fix z."""
    z = np.array([0.0, 0.0, 1.0])

    def f(q):
        x = cartesian(lat, q)
        M = np.einsum("ni,j->nij", x, z) + np.einsum("i,nj->nij", z, x)
        return direct_operator(lat, M)
    return f


def _complex_rows(vec, phase):
    """Complex rows [10, v] * phase: the large scalar keeps every Hermitian overlap between
    vertices positive, so the antiphases (src/utilities.tpp:567-577) are 1 and the complex
    interpolation of a linear field is exact as well."""
    rows = np.hstack([np.full((vec.shape[0], 1), 10.0), vec]) * phase
    return rows[:, None, :]


def _points(n=3000, seed=1):
    return np.random.default_rng(seed).uniform(-3, 3, size=(n, 3))


@pytest.mark.parametrize("case", [HEX, ORTHO])
def test_real_lattice_vectors_rotate_by_Ridx(case):
    """rip_real (src/interpolator_real.tpp:41-45): v -> W[Ridx] v, for real values and complex vectors."""
    lat, bz, g = field_grid(*case)
    f = _vector_field(lat)
    rlu = np.asarray(g.rlu)
    t = fill_field(g, f(rlu), [0, 3, 0, VECTOR, REAL_LATTICE], _complex_rows(f(rlu), 1.0 + 0.5j),
                   [1, 3, 0, VECTOR, REAL_LATTICE])
    Q = _points()
    v, c, aux = _run(t, Q)
    W = _w(t)
    want = np.einsum("nij,nj->ni", W[aux["rot"]], f(aux["ir_q"]))
    assert np.max(np.abs(v - want)) < 1e-10
    assert np.max(np.abs(c[:, 1:] - (1.0 + 0.5j) * want)) < 1e-10
    other = np.einsum("nij,nj->ni", W[aux["invrot"]], f(aux["ir_q"]))
    field = f(Q - aux["tau"])
    if case is ORTHO:
        # operations are their own inverses: the result is the field itself
        assert np.max(np.abs(v - field)) < 1e-10
    else:
        # 3- and 6-fold operations: W[Ridx] and W[invRidx] differ, and the formula tells them apart
        assert np.max(np.abs(want - other)) > 0.1
        assert np.max(np.abs(v - field)) > 0.1


def test_axial_vectors_take_the_determinant_and_invRidx():
    """rip_axial (src/interpolator_axial.tpp:45-50): v -> det(W[Ridx]) W[invRidx] v.  In 6mm the
    field J x is axial (mirrors reverse it), so the result is the field at Q - tau: a wrong
    sign, matrix or transpose fails."""
    lat, bz, g = field_grid(*HEX_POLAR)
    f = _axial_field(lat)
    rlu = np.asarray(g.rlu)
    t = fill_field(g, f(rlu), [0, 3, 0, PSEUDOVECTOR, REAL_LATTICE], _complex_rows(f(rlu), 2.0 - 1.0j),
                   [1, 3, 0, PSEUDOVECTOR, REAL_LATTICE])
    Q = _points(seed=2)
    v, c, aux = _run(t, Q)
    field = f(Q - aux["tau"])
    assert np.max(np.abs(field)) > 0.1
    assert np.max(np.abs(v - field)) < 1e-10
    assert np.max(np.abs(c[:, 1:] - (2.0 - 1.0j) * field)) < 1e-10
    W = _w(t)
    det = np.round(np.linalg.det(W))
    assert (det[aux["rot"]] < 0).any() and (det[aux["rot"]] > 0).any()
    want = det[aux["rot"]][:, None] * np.einsum("nij,nj->ni", W[aux["invrot"]], f(aux["ir_q"]))
    assert np.max(np.abs(v - want)) < 1e-10


def test_axial_vector_formula_in_a_centrosymmetric_group():
    """The same formula with the proper-vector field in 6/mmm (no linear axial field exists there)."""
    lat, bz, g = field_grid(*HEX)
    f = _vector_field(lat)
    t = fill_field(g, f(np.asarray(g.rlu)), [0, 3, 0, PSEUDOVECTOR, REAL_LATTICE])
    Q = _points(seed=3)
    v, c, aux = _run(t, Q)
    W = _w(t)
    det = np.round(np.linalg.det(W))
    want = det[aux["rot"]][:, None] * np.einsum("nij,nj->ni", W[aux["invrot"]], f(aux["ir_q"]))
    assert np.max(np.abs(v - want)) < 1e-10


@pytest.mark.parametrize("case", [HEX_POLAR, ORTHO_POLAR])
def test_real_lattice_matrices(case):
    """rip_real matrices (src/interpolator_real.tpp:47-54): M -> W[Ridx] M W[invRidx]."""
    lat, bz, g = field_grid(*case)
    f = _operator_field(lat)
    t = fill_field(g, f(np.asarray(g.rlu)), [0, 0, 9, VECTOR, REAL_LATTICE])
    Q = _points(seed=4)
    v, c, aux = _run(t, Q)
    W = _w(t)
    M0 = f(aux["ir_q"]).reshape(-1, 3, 3)
    want = (W[aux["rot"]] @ M0 @ W[aux["invrot"]]).reshape(-1, 9)
    assert np.max(np.abs(v - want)) < 1e-10
    field = f(Q - aux["tau"])
    if case is ORTHO_POLAR:
        assert np.max(np.abs(v - field)) < 1e-10
    else:
        assert np.max(np.abs(v - field)) > 0.1


def test_axial_matrices():
    """rip_axial matrices (src/interpolator_axial.tpp:52-58): M -> W[invRidx] M W[Ridx], no
    determinant; for the operator field x z^T + z x^T in 6mm this is the field at Q - tau."""
    lat, bz, g = field_grid(*HEX_POLAR)
    f = _operator_field(lat)
    t = fill_field(g, f(np.asarray(g.rlu)), [0, 0, 9, PSEUDOVECTOR, REAL_LATTICE])
    Q = _points(seed=5)
    v, c, aux = _run(t, Q)
    assert np.max(np.abs(v - f(Q - aux["tau"]))) < 1e-10
    W = _w(t)
    M0 = f(aux["ir_q"]).reshape(-1, 3, 3)
    assert np.max(np.abs(v - (W[aux["invrot"]] @ M0 @ W[aux["rot"]]).reshape(-1, 9))) < 1e-10


def test_mixed_scalars_vectors_matrices_in_one_branch():
    """A branch of 1 scalar + 2 vectors + 1 matrix (the element order of
    count_scalars_vectors_matrices, src/interpolator.hpp:309-332): each part rotated, the scalar
    untouched."""
    lat, bz, g = field_grid(*HEX_POLAR)
    fv, fm = _vector_field(lat), _operator_field(lat)
    rlu = np.asarray(g.rlu)
    vals = np.hstack([np.full((rlu.shape[0], 1), 1.5), fv(rlu), 2 * fv(rlu), fm(rlu)])
    t = fill_field(g, vals, [1, 6, 9, VECTOR, REAL_LATTICE])
    Q = _points(600, seed=6)
    v, c, aux = _run(t, Q)
    W = _w(t)
    Wr, Wi = W[aux["rot"]], W[aux["invrot"]]
    vq, mq = fv(aux["ir_q"]), fm(aux["ir_q"]).reshape(-1, 3, 3)
    assert np.allclose(v[:, 0], 1.5, rtol=0, atol=1e-12)
    assert np.max(np.abs(v[:, 1:4] - np.einsum("nij,nj->ni", Wr, vq))) < 1e-10
    assert np.max(np.abs(v[:, 4:7] - 2 * np.einsum("nij,nj->ni", Wr, vq))) < 1e-10
    assert np.max(np.abs(v[:, 7:] - (Wr @ mq @ Wi).reshape(-1, 9))) < 1e-10
