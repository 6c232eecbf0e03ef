"""The oracle is pinned before anything is compared against it.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails: golden vectors from
the reference's own fixture, closed forms, invariants and brute force.
"""
import numpy as np
import pytest

from conftest import linear_field_grid, nacl_grid, phase_free_vectors
from oracle import oracle as O


def test_nacl_golden_values_and_vectors(nacl):
    """wrap/tests/test_5_gamma.py:95-178 on the oracle: the 48 symmetry-equivalent q_nu must give
    the Euphonic eigenvalues and, up to a per-mode phase, the Euphonic eigenvectors."""
    d = nacl
    lat, bz, g, perm, q_nu, _ = nacl_grid(d)
    assert np.allclose(d["grid_rlu"][perm], g.rlu)
    rc, vals, vecs, aux, _ = O.ir_interpolate_at(g.tables(), q_nu)
    assert rc == 0
    v = np.squeeze(vals)
    assert not np.any(np.isclose(np.diff(v, axis=1), 0.0))
    assert np.allclose(np.diff(v, axis=0), 0.0)
    assert np.allclose(v, d["euphonic_values"])
    assert np.allclose(v, d["brille_values"], rtol=0, atol=1e-9)
    br = np.einsum("ba,ijkb->ijka", d["basis_vectors"], vecs)
    assert np.allclose(phase_free_vectors(br, d["euphonic_vectors"]), d["euphonic_vectors"])
    assert np.allclose(phase_free_vectors(vecs, d["brille_vectors"]), d["brille_vectors"], atol=1e-9)


def test_nacl_golden_cartesian_vectors(nacl):
    """wrap/tests/test_5_gamma.py:180-221: Gamma rotation of Cartesian (angstrom) eigenvectors."""
    d = nacl
    lat, bz, g, perm, q_nu, _ = nacl_grid(d)
    cart = np.einsum("ba,ijkb->ijka", d["basis_vectors"], d["grid_vectors"])
    g.fill(d["grid_values"][perm], d["grid_values_elements"].astype(int), d["grid_values_weights"],
           cart[perm], np.array([0, 24, 0, 2, 1, 0, 0]), d["grid_vectors_weights"], False)
    rc, vals, vecs, aux, _ = O.ir_interpolate_at(g.tables(), d["q_nu"])
    assert rc == 0
    assert np.allclose(phase_free_vectors(vecs, d["euphonic_vectors"]), d["euphonic_vectors"])


def test_nacl_rotation_tables_move_every_q_nu_to_the_grid_point(nacl):
    """Stage (a) alone: every q_nu = R^T q_ir folds back onto q_ir, with R^T_invR q = ir_q and
    R_Ridx the inverse of R_invRidx."""
    d = nacl
    lat, bz, g, perm, q_nu, _ = nacl_grid(d)
    t = g.tables()
    rc, q, tau, r, ir, st = O.ir_moveinto(t, q_nu)
    assert rc == 0 and not st.any()
    assert np.allclose(q, d["grid_rlu"][4], atol=1e-12)
    Rs = t["rotations"].reshape(-1, 3, 3)
    for k in range(len(q_nu)):
        assert np.array_equal(Rs[r[k]] @ Rs[ir[k]], np.eye(3, dtype=int))
        assert np.allclose(Rs[ir[k]].T @ (q_nu[k] - tau[k]), q[k], atol=1e-12)


def test_linear_field_is_reproduced_in_the_irreducible_zone():
    """src/tests/trellis_test.cpp:76-137: the vertex xyz stored as data is interpolated back
    to the point's xyz within 2e-10 (linear interpolation is exact for linear fields)."""
    lat, bz, g, = linear_field_grid()
    t = g.tables()
    rng = np.random.default_rng(5)
    ir = np.asarray(bz.ir_vertices)
    w = rng.dirichlet(np.ones(len(ir)), size=400)
    q = w @ ir  # random points inside the convex irreducible polyhedron
    rc, vals, _, aux, _ = O.ir_interpolate_at(t, q, no_move=True)
    assert rc == 0
    xyz = q @ np.asarray(lat.reciprocal_vectors).T
    assert np.max(np.abs(vals[:, 0, :] - xyz)) < 2e-10


def test_linear_field_everywhere_with_reciprocal_rotation():
    """For a cubic lattice the xyz field rotates like a reciprocal-lattice vector, so after
    fold-in, interpolation and rip_recip the result is xyz(Q - tau) for any Q."""
    lat, bz, g = linear_field_grid(3e-4)
    t = g.tables()
    Q = np.random.default_rng(6).uniform(-3, 3, size=(3000, 3))
    rc, vals, _, aux, _ = O.ir_interpolate_at(t, Q)
    assert rc == 0
    B = np.asarray(lat.reciprocal_vectors)
    q1 = Q - aux["tau"]
    assert np.max(np.abs(vals[:, 0, :] - q1 @ B.T)) < 2e-10


def test_weights_are_barycentric():
    """Sum of weights is one, weights are non-negative within tolerance and reproduce x."""
    lat, bz, g = linear_field_grid(3e-4)
    t = g.tables()
    ir = np.asarray(bz.ir_vertices)
    q = np.random.default_rng(7).dirichlet(np.ones(len(ir)), size=500) @ ir
    x = q @ np.asarray(lat.reciprocal_vectors).T
    rc, node, tet, niw, idx, w = O.locate(t, x)
    assert rc == 0
    V = t["vertices"]
    for i in range(len(x)):
        n = niw[i]
        assert 1 <= n <= 8
        assert abs(w[i, :n].sum() - 1) < 1e-12
        assert (w[i, :n] > -1e-10).all()
        assert np.allclose((w[i, :n, None] * V[idx[i, :n]]).sum(0), x[i], atol=1e-12)


def test_moveinto_invariants_and_brute_force_tau():
    """Q = q + tau with integer tau and q inside the first zone; for points strictly inside, tau
    is the nearest lattice vector (brute force over the neighbourhood)
    (src/tests/brillouinzone_test.cpp:114-150)."""
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    Q = S.q_points(1, 2000, stream=11)
    rc, q, tau, st = O.moveinto(t, Q)
    assert rc == 0
    assert np.allclose(q + tau, Q, atol=1e-12, rtol=0)
    B = t["outer_to_xyz"].reshape(3, 3)
    # brute force: the lattice vector nearest Q in the Cartesian metric
    cand = np.array([[i, j, k] for i in range(-6, 7) for j in range(-6, 7) for k in range(-6, 7)])
    for i in range(0, 2000, 7):
        d = np.linalg.norm((Q[i] - cand) @ B.T, axis=1)
        o = np.argsort(d)
        if d[o[1]] - d[o[0]] > 1e-9:
            assert np.array_equal(cand[o[0]], tau[i])


def test_wedge_choice_is_unique_for_generic_points_and_first_for_ties():
    """Brute force over the point group: a generic q has exactly one R with R^T q in the wedge,
    and ir_moveinto returns the first such R in PointSymmetry order."""
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    Q = S.q_points(1, 500, stream=12)
    rc, q1, tau, st = O.moveinto(t, Q)
    rc2, q, tau2, r, ir, st2 = O.ir_moveinto(t, Q)
    assert rc == 0 and rc2 == 0
    Rs = t["rotations"].reshape(-1, 3, 3)
    G = t["outer_metric_recip"].reshape(3, 3)
    n = t["wedge_normals"]
    for k in range(len(Q)):
        inside = []
        for j in range(len(Rs)):
            p = Rs[j].T @ q1[k]
            dots = n @ G @ p
            inside.append(bool((dots >= -1e-12).all()))
        if np.count_nonzero(inside) == 1:
            assert inside.index(True) == ir[k] or (ir[k] == t["identity_index"] and inside[t["identity_index"]])
        assert inside[ir[k]]


def test_primitives_against_closed_forms():
    """orient3d against the determinant formula (six times the signed volume), the lattice form
    against the Cartesian one, find_bin against bisection, antiphase against the polar angle."""
    rng = np.random.default_rng(8)
    for _ in range(50):
        a, b, c, d = rng.normal(size=(4, 3))
        det = np.linalg.det(np.array([a - d, b - d, c - d]))
        assert np.isclose(O.orient3d_xyz(a, b, c, d), det, rtol=1e-12, atol=1e-14)
    from brille_amd import synthetic as S
    lat = S.lattice(1)
    B = np.asarray(lat.reciprocal_vectors)
    Gr = B.T @ B
    A = np.asarray(lat.real_vectors)
    Gd = A.T @ A
    scale = np.linalg.det(B) / (2 * np.pi)
    for _ in range(50):
        h = rng.normal(size=(4, 3))
        lat_val = O.orient3d_lattice(*h, Gr.ravel(), Gd.ravel(), scale)
        x = h @ B.T
        assert np.isclose(lat_val, np.linalg.det(np.array([x[0] - x[3], x[1] - x[3], x[2] - x[3]])), rtol=1e-10)
    edges = np.cumsum(rng.uniform(0.1, 1, 20))
    import bisect
    for x in rng.uniform(edges[0], edges[-1], 200):
        assert O.find_bin(edges, x) == bisect.bisect_right(edges, x) - 1
    assert O.find_bin(edges, edges[0] - 1) == 0
    for _ in range(20):
        u = rng.normal(size=6) + 1j * rng.normal(size=6)
        v = rng.normal(size=6) + 1j * rng.normal(size=6)
        z = np.vdot(u, v)
        assert np.isclose(O.antiphase(u, v), np.exp(-1j * np.angle(z)))


def test_approx_matches_the_reference_tolerance_formula():
    """approx_float::_scalar: |a-b| <= Tabs + Trel |a+b| with Trel = max(eps*tol*1e4, Ttol)."""
    eps = np.finfo(float).eps
    assert O.approx(1.0, 1.0 + 1.5e-12)
    assert not O.approx(1.0, 1.0 + 1e-11)
    assert O.approx(0.0, 4e-15)
    assert not O.approx(0.0, 6e-15)
    assert O.approx(0.0, 9e-13, 1e-12, 1000)
    assert not O.approx(0.0, 1.1e-12, 1e-12, 1000)
    assert O.approx(1.0, 1.0 + 2.0 * 1e7 * eps * 0.99, 1e-12, 1000)


def test_error_statuses():
    """Non-finite Q cannot be folded in (the reference throws 'Not all points inside Brillouin
    zone'); an unfilled grid is reported before any work."""
    from brille_amd import synthetic as S
    g = S.grid(0)
    t = g.tables()
    Q = np.array([[0.1, 0.2, 0.3], [np.nan, 0, 0], [0.2, 0.1, 0.0]])
    rc, vals, vecs, aux, bad = O.ir_interpolate_at(t, Q)
    assert rc == 3 and bad == 1
    assert list(aux["status"]) == [0, 1, 0]
    rc, *_ = O.ir_interpolate_at(t, np.zeros((0, 3)))
    assert rc == 0


def test_permutation_table_is_applied():
    """PermutationTable::safe_get: a permutation stored for (pivot, j) reorders j's modes."""
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    Q = S.q_points(1, 50, stream=13)
    rc, v0, c0, aux, _ = O.ir_interpolate_at(t, Q)
    i0, i1 = int(aux["iw_index"][0, 0]), int(aux["iw_index"][0, 1])
    V = t["vertices"].shape[0]
    B = int(t["branches"])
    perm = np.arange(B)[::-1].copy()
    # swapping the stored modes of vertex i1 and undoing it with a permutation leaves results unchanged
    vals = t["values"].copy()
    vecs = t["vectors"].copy()
    vals[i1] = vals[i1][perm]
    vecs[i1] = vecs[i1][perm]
    keys = np.array([i0 * V + i1], np.uint64)
    rc, v1, c1, aux1, _ = O.ir_interpolate_at(t, Q[:1], values=vals, vectors=vecs,
                                              perm=(keys, np.array([0], np.uint32), perm[None, :].astype(np.uint32)))
    assert rc == 0
    assert np.allclose(v1[0], v0[0], rtol=1e-13, atol=1e-13)
    assert np.allclose(c1[0], c0[0], rtol=1e-12, atol=1e-12)


def _first_zone_points(g, n, seed):
    """Points whose first-zone image is inside the irreducible wedge: random
    interior points of the irreducible polyhedron, shifted by lattice vectors."""
    bz = g.BrillouinZone
    ir = np.asarray(bz.ir_vertices)
    r = np.random.default_rng(seed)
    w = r.dirichlet(np.ones(len(ir)), size=n)
    q = w @ ir
    tau = r.integers(-2, 3, size=(n, 3))
    return q, q + tau


def test_first_zone_mode_matches_the_irreducible_path_on_wedge_images():
    """interpolate_at (moveinto, no rotation) equals ir_interpolate_at wherever the
    wedge search picks the identity, and its q is BrillouinZone::moveinto's."""
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    q, Q = _first_zone_points(g, 3000, 5)
    rc1, v1, c1, a1, _ = O.ir_interpolate_at(t, Q, first_zone=True)
    rc0, v0, c0, a0, _ = O.ir_interpolate_at(t, Q)
    assert rc0 == 0 and rc1 == 0
    assert np.all(a0["rot"] == t["identity_index"]) and np.all(a1["rot"] == t["identity_index"])
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(c1, c0)
    rc, qm, taum, st = O.moveinto(t, Q)
    np.testing.assert_array_equal(a1["ir_q"], qm)
    np.testing.assert_array_equal(a1["tau"], taum)
    np.testing.assert_allclose(a1["ir_q"], q, atol=1e-12)


def test_first_zone_mode_fails_outside_the_wedge():
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    Q = S.q_points(1, 2000, stream=41)
    rc, v, c, aux, bad = O.ir_interpolate_at(t, Q, first_zone=True)
    assert rc == 5  # BRILLE_ERR_NULL_NODE: first-zone images outside the trellis
    assert aux["status"][bad] == 3  # BRILLE_PT_NULL_NODE


def test_earlier_stage_errors_win_over_earlier_indices():
    """The reference folds every point before it locates any (src/bz_trellis.hpp:164-168) and
    checks moveinto before the wedge (src/bz_move.cpp:158, 288-294): a stage-(a) failure at a
    later index is the error even when a stage-(b) failure comes first in the input."""
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    Qb = S.q_points(1, 3000, stream=42)
    rc, *_, bad = O.ir_interpolate_at(t, Qb, first_zone=True)
    assert rc == 5  # first-zone images outside the wedge: null nodes
    Q = np.vstack([Qb[bad:bad + 1], Qb[:50], [[np.nan, 0.0, 0.0]], Qb[:5]])
    rc, v, c, aux, bad2 = O.ir_interpolate_at(t, Q, first_zone=True)
    assert aux["status"][0] == 3 and aux["status"][51] == 1
    assert rc == 3 and bad2 == 51  # BRILLE_ERR_OUTSIDE_BZ at the NaN, not the null node at 0


def test_table_only_locate_equals_the_full_path():
    """ir_moveinto + get_xyz + locate (tables only, no payload) give the full path's q, tau, Ridx,
    node, tetrahedron and weights bit for bit: how a payload too large for the host is checked."""
    from brille_amd import synthetic as S
    g = S.grid(1)
    t = g.tables()
    Q = S.q_points(1, 20000, stream=44)
    rc, v, c, aux, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    rc1, q, tau, r, ir, st = O.ir_moveinto(t, Q)
    rc2, node, tet, niw, idx, w = O.locate(t, O.get_xyz(t, q))
    assert rc == rc1 == rc2 == 0
    for a, b in ((q, aux["ir_q"]), (tau, aux["tau"]), (r, aux["rot"]), (node, aux["node"]), (tet, aux["tet"]),
                 (niw, aux["n_iw"]), (idx, aux["iw_index"]), (w, aux["iw_weight"])):
        np.testing.assert_array_equal(a, b)
