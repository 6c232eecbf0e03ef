"""Host-side constructors: space groups, lattices, zones and trellises (CPU only)."""
import numpy as np
import pytest

import brille_amd as b


def _pg_order(n):
    bounds = [(1, 1), (2, 2), (9, 2), (15, 4), (24, 4), (46, 4), (74, 8), (80, 4), (82, 4), (88, 8), (98, 8),
              (110, 8), (122, 8), (142, 16), (146, 3), (148, 6), (155, 6), (161, 6), (167, 12), (173, 6), (174, 6),
              (176, 12), (182, 12), (186, 12), (190, 12), (194, 24), (199, 12), (206, 24), (214, 24), (220, 24),
              (230, 48)]
    for hi, o in bounds:
        if n <= hi:
            return o


@pytest.mark.parametrize("n", range(1, 231))
def test_every_space_group_generates_with_its_order(n):
    s = b.Symmetry(str(n))
    W = s.W
    centring = sum(1 for k in range(len(s)) if np.array_equal(W[k], np.eye(3, dtype=int)))
    assert len(s) == _pg_order(n) * centring
    # closure: the product of any two operations is in the group (mod lattice translations)
    ops = [(W[k], s.w[k]) for k in range(len(s))]
    for i in range(0, len(ops), max(1, len(ops) // 6)):
        for j in range(0, len(ops), max(1, len(ops) // 6)):
            R = ops[i][0] @ ops[j][0]
            t = (ops[i][0] @ ops[j][1] + ops[i][1]) % 1.0
            assert any(np.array_equal(R, o[0]) and np.allclose((t - o[1] + 0.5) % 1.0 - 0.5, 0, atol=1e-9)
                       for o in ops)


def test_reciprocal_basis_is_dual():
    lat = b.Direct(np.array([5.1, 7.3, 9.0]), np.array([90, 104, 90]), "14")
    A = np.asarray(lat.real_vectors)
    B = np.asarray(lat.reciprocal_vectors)
    assert np.allclose(A.T @ B, 2 * np.pi * np.eye(3))
    assert np.isclose(lat.a, 5.1) and np.isclose(lat.b, 7.3) and np.isclose(lat.c, 9.0)
    assert np.isclose(np.degrees(lat.beta), 104)


@pytest.mark.parametrize("args,order", [
    (((4.96,) * 3, (90,) * 3, "227"), 48),
    (((3.2, 3.2, 5.2), (90, 90, 120), "194"), 24),
    (((3.0, 3.0, 7.0), (90, 90, 90), "139"), 16),
    (((5.0, 6.0, 7.0), (90, 90, 90), "62"), 8),
    (((5.1, 7.3, 9.0), (90, 104, 90), "14"), 4),
    (((4.9, 4.9, 5.4), (90, 90, 120), "154"), 6),
    (((3.0, 4.0, 5.0), (80, 95, 105), "1"), 1),
    (((3.0, 4.0, 5.0), (80, 95, 105), "2"), 2),
])
def test_zone_volumes(args, order):
    lat = b.Direct(np.array(args[0], float), np.array(args[1], float), args[2])
    bz = b.BrillouinZone(lat)
    conv_vol = abs(np.linalg.det(np.asarray(lat.reciprocal_vectors)))
    cent = {1: 1, 2: 2, 3: 2, 4: 2, 5: 2, 6: 4, 7: 3}[lat.bravais]
    # the primitive reciprocal cell (= first zone volume) is `cent` times the conventional one
    assert np.isclose(bz.polyhedron.volume, conv_vol * cent, rtol=1e-10)
    assert np.isclose(bz.ir_polyhedron.volume * order, bz.polyhedron.volume, rtol=1e-9)


def test_time_reversal_halves_a_polar_group():
    lat = b.Direct(np.array([4.9, 4.9, 5.4]), np.array([90, 90, 120.0]), "154")
    v0 = b.BrillouinZone(lat).ir_polyhedron.volume
    v1 = b.BrillouinZone(lat, time_reversal_symmetry=True).ir_polyhedron.volume
    assert np.isclose(v0, 2 * v1)


@pytest.mark.parametrize("cfg,vf", [(0, None), (1, None), (4, 1e-4)])
def test_trellis_invariants(cfg, vf):
    """Node volumes sum to the irreducible volume (src/tests/trellis_test.cpp:291-341) and Gamma
    is a vertex exactly once (:254-262); cube nodes are axis-aligned boxes; tetrahedra are
    positively oriented."""
    from brille_amd import synthetic as S
    g = S.grid(cfg, volume_fraction=vf)
    t = g.tables()
    assert np.isclose(g.total_node_volume(), g.BrillouinZone.ir_polyhedron.volume, rtol=1e-10)
    V = t["vertices"]
    assert np.count_nonzero(np.all(np.abs(V) < 1e-12, axis=1)) == 1
    cv = V[t["cube_vertices"]]
    assert np.allclose(cv[:, 0] + cv[:, 7], cv[:, 1] + cv[:, 6])
    tv = V[t["tet_vertices"]]
    vol = np.einsum("ij,ij->i", tv[:, 0] - tv[:, 3], np.cross(tv[:, 1] - tv[:, 3], tv[:, 2] - tv[:, 3])) / 6
    assert (vol > 0).all()
    assert np.allclose(vol, t["tet_volume"], rtol=1e-9)
    c = t["tet_circumsphere"]
    d = np.linalg.norm(tv - c[:, None, :3], axis=2)
    assert np.allclose(d, c[:, 3:4], rtol=1e-8)
    # every non-null node is cube or poly and indexes a valid record
    nt = t["node_type"]
    assert t["node_index"][nt == 3].max() < t["cube_vertices"].shape[0]
    assert t["node_index"][nt == 4].max() + 1 == t["poly_tet_offset"].shape[0] - 1


def test_nacl_grid_matches_the_reference_fixture(nacl):
    """The reference's NaCl grid vertices (test_5_gamma.npz grid_rlu) are reproduced exactly,
    up to order (wrap/tests/test_5_gamma.py:97-105)."""
    from conftest import nacl_grid
    lat, bz, g, perm, q_nu, _ = nacl_grid(nacl)
    assert len(perm) == len(nacl["grid_rlu"]) == 29
    assert np.allclose(nacl["grid_rlu"][perm], g.rlu, atol=1e-12)


def test_fill_validation():
    from brille_amd import synthetic as S
    g = S.grid(0)
    V = g.rlu.shape[0]
    with pytest.raises(RuntimeError, match="arrays but"):
        g.fill(np.zeros((V + 1, 6, 1)), np.array([1, 0, 0]), np.zeros((V + 1, 6, 1), complex), np.array([1, 0, 0]))
    with pytest.raises(Exception, match="3N"):
        g.fill(np.zeros((V, 6, 2)), np.array([0, 2, 0]), np.zeros((V, 6, 1), complex), np.array([1, 0, 0]))
    with pytest.raises(RuntimeError, match="RotatesLike"):
        g.fill(np.zeros((V, 6, 1)), np.array([1, 0, 0, 7]), np.zeros((V, 6, 1), complex), np.array([1, 0, 0]))
    # sort=True is accepted; the sorting itself runs on the device at first use
    g.fill(np.zeros((V, 6, 1)), np.array([1, 0, 0]), np.zeros((V, 6, 1), complex), np.array([1, 0, 0]), True)


def test_fill_keeps_references():
    """wrap/tests/test_4_interpolation.py:265-277: fill holds the caller's arrays."""
    import sys
    from brille_amd import synthetic as S
    g = S.grid(0)
    V = g.rlu.shape[0]
    vals = np.zeros((V, 6, 1))
    vecs = np.zeros((V, 6, 1), complex)
    n0 = sys.getrefcount(vals)
    g.fill(vals, np.array([1, 0, 0]), vecs, np.array([1, 0, 0]))
    assert sys.getrefcount(vals) == n0 + 1
    assert g.values is vals
