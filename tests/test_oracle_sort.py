"""The oracle's mode sort (DualInterpolator::sort, src/interpolatordual.hpp:398-434)
pinned against closed forms and brute force (CPU only).

* the assignment solver against exhaustive search over all permutations of
  random matrices (unique optimum) and on matrices whose optimum is a known
  permutation;
* every cost function on vectors whose angles, distances and products are
  known in closed form (src/interpolator.hpp:246-305, src/utilities.tpp:249-417),
  with the antiphase making a global phase irrelevant
  (src/interpolator_cost.tpp:26-41);
* the key set (PolyTrellis::collect_keys) against the vertex sets the oracle's
  own location returns for the centre of every cube and tetrahedron;
* the whole table on a payload whose modes were shuffled at every vertex:
  the sort must find, for every key (i, j), the mode of j that carries the
  same original mode as each mode of i.
"""
import itertools

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7])
def test_assignment_is_the_exhaustive_optimum(n):
    r = np.random.default_rng(100 + n)
    perms = list(itertools.permutations(range(n)))
    for _ in range(40):
        C = r.random((n, n))
        totals = [sum(C[i, p[i]] for i in range(n)) for p in perms]
        best = perms[int(np.argmin(totals))]
        row, col = O.lap(C)
        assert tuple(row) == best
        assert all(col[row[i]] == i for i in range(n))


def test_assignment_recovers_a_planted_permutation():
    r = np.random.default_rng(7)
    for n in (8, 16, 40):
        p = r.permutation(n)
        C = r.uniform(1.0, 2.0, size=(n, n))
        C[np.arange(n), p] = r.uniform(0.0, 0.1, size=n)
        row, col = O.lap(C)
        assert np.array_equal(row, p)
        assert np.array_equal(col, np.argsort(p))


def _two_vertex_tables(B, sv_el, sc_el):
    g = S.grid(1)
    t = dict(g.tables())
    t["branches"] = B
    t["values_elements"] = np.array(sv_el)
    t["vectors_elements"] = np.array(sc_el)
    return t


def _cost(t, vals, vecs, vcost=None, ccost=None):
    V = np.asarray(t["vertices"]).shape[0]
    B = int(t["branches"])
    values = np.zeros((V, B, int(np.sum(t["values_elements"]))))
    vectors = np.zeros((V, B, int(np.sum(t["vectors_elements"]))), complex)
    values[:2] = vals
    vectors[:2] = vecs
    rc, row, col, cost = O.sort_pairs(t, np.array([[0, 1]]), values, vectors, vcost, ccost, want_cost=True)
    assert rc == 0
    return cost[0], row[0], col[0]


def test_default_costs_in_closed_form():
    t = _two_vertex_tables(2, [1, 0, 0], [0, 3, 0])
    e1, e2 = np.eye(3)[0], np.eye(3)[1]
    vals = np.array([[[1.0], [5.0]], [[2.0], [3.0]]])
    vecs = np.array([[e1, e2], [np.exp(0.7j) * e2, np.exp(-1.1j) * (e1 + e2) / np.sqrt(2)]])
    cost, row, col = _cost(t, vals, vecs)
    # |value differences| + sin^2 of the Hermitian angles (1, 1/2, 0, 1/2)
    want = np.array([[1 + 1, 2 + 0.5], [3 + 0, 2 + 0.5]])
    np.testing.assert_allclose(cost, want, rtol=0, atol=1e-14)
    assert list(row) == [0, 1] and list(col) == [0, 1]


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_vector_cost_kinds(kind):
    t = _two_vertex_tables(2, [1, 0, 0], [0, 3, 0])
    th = 0.4
    a0 = np.array([1.0, 0, 0])
    a1 = np.array([0, 1.0, 0])
    b0 = np.exp(0.3j) * np.array([np.cos(th), np.sin(th), 0])  # Hermitian angle th to a0 once phased
    b1 = np.exp(-2.0j) * np.array([0, 0, 1.0])
    vals = np.zeros((2, 2, 1))
    vecs = np.array([[a0, a1], [b0, b1]])
    w = (0.0, 1.0, 0.0)
    cost, _, _ = _cost(t, vals, vecs, capi.cost_info(0, 0, w), capi.cost_info(0, kind, w))
    # after the antiphase the phased b equals its real form when the overlap is non-zero
    if kind == 1:  # vector distance
        want00 = np.linalg.norm(a0 - np.array([np.cos(th), np.sin(th), 0]))
    elif kind == 2:  # 1 - |a^H b|^2
        want00 = 1 - np.cos(th) ** 2
    else:  # Euclidean (3) and Hermitian (4) angles agree once phased
        want00 = th
    np.testing.assert_allclose(cost[0, 0], want00, rtol=0, atol=1e-14)
    # a0 against b1: orthogonal (angle pi/2, product 0, distance sqrt 2)
    want01 = {1: np.sqrt(2), 2: 1.0, 3: np.pi / 2, 4: np.pi / 2}[kind]
    np.testing.assert_allclose(cost[0, 1], want01, rtol=0, atol=1e-14)


def test_weights_scalar_and_matrix_terms():
    t = _two_vertex_tables(1, [1, 3, 9], [1, 0, 9])
    r = np.random.default_rng(3)
    va, vb = r.normal(size=13), r.normal(size=13)
    ca, cb = r.normal(size=10) + 1j * r.normal(size=10), r.normal(size=10) + 1j * r.normal(size=10)
    vals = np.array([[va], [vb]])
    vecs = np.array([[ca], [cb]])
    vw, cw = (2.0, 3.0, 5.0), (7.0, 0.0, 11.0)
    cost, _, _ = _cost(t, vals, vecs, capi.cost_info(0, 1, vw), capi.cost_info(0, 1, cw))
    s = abs(va[0] - vb[0])
    v = np.linalg.norm(va[1:4] - vb[1:4])
    m = np.linalg.norm(va[4:] - vb[4:])
    h = np.vdot(ca, cb)
    ph = np.exp(-1j * np.angle(h)) * cb
    cs = abs(ca[0] - ph[0])
    cm = np.linalg.norm(ca[1:] - ph[1:])
    want = (vw[0] * s + vw[1] * v + vw[2] * m) + (cw[0] * cs + cw[2] * cm)
    np.testing.assert_allclose(cost[0, 0], want, rtol=1e-13)


@pytest.mark.parametrize("cfg", [0, 1])
def test_keys_are_the_vertex_sets_of_cubes_and_tetrahedra(cfg):
    t = S.grid(cfg).tables()
    keys = O.sort_keys(t)
    x = np.asarray(t["vertices"])
    centres = [x[np.asarray(t["cube_vertices"]).reshape(-1, 8)].mean(axis=1),
               x[np.asarray(t["tet_vertices"]).reshape(-1, 4)].mean(axis=1)]
    rc, node, tet, niw, idx, w = O.locate(t, np.vstack(centres))
    assert rc == 0
    V = x.shape[0]
    got = set()
    for k in range(len(niw)):
        vs = idx[k, :niw[k]]
        for a in vs:
            for b in vs:
                if a != b:
                    got.add(int(a) * V + int(b))
    assert got == set(int(k) for k in keys)


@pytest.mark.parametrize("cfg", [0, 1])
def test_sort_undoes_shuffled_modes(cfg):
    t = dict(S.grid(cfg).tables())
    values, vectors, order = S.shuffled_modes(t, seed=cfg)
    rc, keys, pairs, row, col = O.sorted_table(t, values, vectors)
    assert rc == 0 and len(pairs) > 100
    i, j = pairs[:, 0], pairs[:, 1]
    # row[k, a] = b with order[j, b] == order[i, a]
    inv_j = np.argsort(order[j], axis=1)
    want_row = np.take_along_axis(inv_j, order[i], axis=1)
    np.testing.assert_array_equal(row, want_row)
    inv_i = np.argsort(order[i], axis=1)
    np.testing.assert_array_equal(col, np.take_along_axis(inv_i, order[j], axis=1))
    # the unshuffled payload sorts to the identity everywhere
    rows = np.arange(order.shape[0])[:, None]
    inv = np.argsort(order, axis=1)
    rc, _, _, row0, col0 = O.sorted_table(t, values[rows, inv], vectors[rows, inv])
    assert rc == 0
    assert (row0 == np.arange(row0.shape[1])).all() and (col0 == np.arange(col0.shape[1])).all()
