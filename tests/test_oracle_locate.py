"""The oracle's point location pinned on hand-built trellises.

PolyTrellis::node_subscript (src/trellis_poly.hpp:382-434) steps to a
neighbouring node when the first-found node is null and the point lies on a
bin boundary; PolyNode::indices_weights (src/trellis_node.hpp:275-308) takes
the first tetrahedron that contains the point and, when none does, the one with
the largest "most negative" score (the circumsphere rejection -d^2, or the
smallest barycentric weight), recomputed without the circumsphere shortcut.

The expectations here come from the construction, not from the oracle's code:
* node steps: brute force over all 64 nodes of a 4 x 4 x 4 trellis for the
  non-null nodes whose closed box holds the point; where exactly one does, the
  oracle must pick it;
* tetrahedra: dyadic vertices, so barycentric weights (numpy solve here,
  orient3d in the oracle) and circumsphere distances are known exactly.
"""
import numpy as np
import pytest

from conftest import CUBE, POLY, hand_trellis, node_id
from oracle import oracle as O

NULL_NODE = 0xFFFFFFFF
TOL = 1e-11


def _locate(t, x):
    x = np.atleast_2d(np.asarray(x, float))
    rc, node, tet, niw, idx, w = O.locate(t, x)
    return rc, node, tet, niw, idx, w


def _containing_non_null(t, x):
    """Brute force: non-null nodes whose closed box (grown by TOL) holds x."""
    out = []
    for k in range(4):
        for j in range(4):
            for i in range(4):
                lo = np.array([i, j, k], float)
                if np.all(x >= lo - TOL) and np.all(x <= lo + 1 + TOL) and t["node_type"][node_id(i, j, k)] in (CUBE, POLY):
                    out.append(node_id(i, j, k))
    return out


CASES = [
    # (node types, point, name): in each, exactly one non-null node's box holds the point
    ({(1, 0, 0): CUBE}, [2.0, 0.5, 0.5], "on a knot plane, step -1"),
    ({(0, 3, 0): CUBE}, [0.5, 3.0 - 1e-13, 0.5], "just below a knot plane, step +1"),
    ({(1, 1, 2): CUBE}, [2.0, 2.0, 2.5], "on an edge, diagonal step (check two)"),
    ({(1, 1, 1): CUBE}, [2.0, 2.0, 2.0], "on a corner, all three steps"),
    ({(3, 1, 1): CUBE}, [4.0, 1.5, 1.5], "on the last knot (find_bin past the end)"),
]


@pytest.mark.parametrize("types,x,name", CASES, ids=[c[2] for c in CASES])
def test_node_steps_find_the_only_usable_node(types, x, name):
    t = hand_trellis(types)
    x = np.array(x)
    want = _containing_non_null(t, x)
    assert len(want) == 1, name
    rc, node, tet, niw, idx, w = _locate(t, x)
    assert rc == 0, name
    assert node[0] == want[0], name
    # the linear field comes back: the located cube's weights reproduce x
    assert abs(w[0, :niw[0]].sum() - 1) < 1e-12
    assert np.allclose((w[0, :niw[0], None] * t["vertices"][idx[0, :niw[0]]]).sum(0), x, atol=1e-12)


def test_first_found_node_wins_when_it_is_usable():
    """Both nodes on either side of x = 2 are cubes: find_bin's node (first knot > x, minus one) is kept."""
    t = hand_trellis({(2, 1, 1): CUBE, (1, 1, 1): CUBE})
    rc, node, *_ = _locate(t, [2.0, 1.5, 1.5])
    assert node[0] == node_id(2, 1, 1)


def test_no_usable_neighbour_is_a_null_node():
    t = hand_trellis({(3, 3, 3): CUBE})
    rc, node, *_ = _locate(t, [2.0, 0.5, 0.5])
    assert rc == 5 and node[0] == NULL_NODE  # BRILLE_ERR_NULL_NODE
    # off every boundary: no step is tried
    rc, node, *_ = _locate(t, [2.5, 2.5, 2.5])
    assert rc == 5


def test_the_pair_loop_also_takes_a_double_step():
    """src/trellis_poly.hpp:405-418 loops j over every direction, j == i included, so before the
    diagonal node it tries two steps along the first boundary direction: with both usable, the
    node two bins away is taken although the point is not in it."""
    t = hand_trellis({(0, 2, 2): CUBE, (1, 1, 2): CUBE})
    rc, node, *_ = _locate(t, [2.0, 2.0, 2.5])
    assert rc == 0
    assert node[0] == node_id(0, 2, 2)
    assert _containing_non_null(t, np.array([2.0, 2.0, 2.5])) == [node_id(1, 1, 2)]


def test_points_below_the_first_knot_use_the_first_bin():
    """find_bin (src/trellis_poly.hpp:66-73) returns bin 0 for x below the first knot."""
    t = hand_trellis({(0, 0, 0): CUBE})
    rc, node, *_ = _locate(t, [-0.25, 0.5, 0.5])
    assert rc == 0 and node[0] == node_id(0, 0, 0)


# ---- PolyNode: tetrahedra of the cube [1, 2]^3 (node (1, 1, 1)) ------------------
A0 = [[1, 1, 1], [2, 1, 1], [1, 2, 1], [1, 1, 2]]      # corner tetrahedron at (1,1,1)
A1 = [[2, 1, 1], [1, 1, 1], [2, 2, 1], [2, 1, 2]]      # its mirror image in x = 1.5
SMALL = [[2, 2, 2], [1.75, 2, 2], [2, 1.75, 2], [2, 2, 1.75]]  # not on the vertex lattice


def positively_oriented(tet):
    """The vertex order whose orient3d is +6V (det[a-d, b-d, c-d] > 0): barycentric weights are
    orient3d / (6V) (src/trellis_node.hpp:326-333), positive inside."""
    p = [list(map(float, v)) for v in tet]
    a, b, c, d = map(np.asarray, p)
    if np.linalg.det(np.array([a - d, b - d, c - d])) < 0:
        p[2], p[3] = p[3], p[2]
    return p


def _poly_tables(tets):
    verts = np.array([[i, j, k] for k in range(5) for j in range(5) for i in range(5)], float)
    extra = [p for tet in tets for p in tet if list(p) not in verts.tolist()]
    t = hand_trellis({(1, 1, 1): POLY}, {(1, 1, 1): [A0]})
    if extra:
        # append the off-lattice vertices and rebuild the tetrahedron tables
        ex = np.unique(np.array(extra, float), axis=0)
        allv = np.vstack([verts, ex])
        t["vertices"] = allv
        t["values"] = allv[:, None, :].copy()
        t["vectors"] = np.zeros((len(allv), 1, 1), complex)
    from conftest import tet_record
    V = t["vertices"].tolist()
    tv, tc, vol = [], [], []
    for tet in tets:
        tet = positively_oriented(tet)
        tv.append([V.index([float(c) for c in p]) for p in tet])
        c, v = tet_record(tet)
        tc.append(c)
        vol.append(v)
    t["tet_vertices"] = np.array(tv, np.uint32)
    t["tet_circumsphere"] = np.array(tc)
    t["tet_volume"] = np.array(vol)
    t["poly_tet_offset"] = np.array([0, len(tets)], np.uint32)
    return t


def _bary(tet, x):
    p = np.asarray(tet, float)
    l = np.linalg.solve((p[1:] - p[0]).T, np.asarray(x, float) - p[0])
    return np.array([1 - l.sum(), *l])


def test_first_containing_tetrahedron_wins():
    """A point on the edge the two tetrahedra share is in both; the first in the node's order wins,
    and only its two non-zero weights are kept."""
    x = [1.5, 1.0, 1.0]
    for order in ([A0, A1], [A1, A0]):
        t = _poly_tables(order)
        rc, node, tet, niw, idx, w = _locate(t, x)
        assert rc == 0 and tet[0] == 0 and niw[0] == 2
        assert np.allclose(np.sort(w[0, :2]), [0.5, 0.5])


def test_fallback_takes_the_largest_most_negative_weight():
    """Outside both tetrahedra (inside both circumspheres): smallest weights -1 (A0) and -0.5 (A1),
    so A1, with all four weights kept (one of them negative)."""
    x = [1.75, 1.625, 1.625]
    assert np.allclose(_bary(A0, x).min(), -1.0) and np.allclose(_bary(A1, x).min(), -0.5)
    t = _poly_tables([A0, A1])
    rc, node, tet, niw, idx, w = _locate(t, x)
    assert rc == 0 and tet[0] == 1 and niw[0] == 4
    got = dict(zip(idx[0, :4].tolist(), w[0, :4].tolist()))
    V = t["vertices"].tolist()
    want = {V.index([float(c) for c in p]): b for p, b in zip(A1, _bary(A1, x))}
    assert got.keys() == want.keys()
    assert all(abs(got[k] - want[k]) < 1e-14 for k in want)


def test_fallback_ties_go_to_the_first_tetrahedron():
    """On the mirror plane x = 1.5 both smallest weights are exactly -0.75: std::max_element keeps
    the first."""
    x = [1.5, 1.625, 1.625]
    assert _bary(A0, x).min() == _bary(A1, x).min() == -0.75
    for order in ([A0, A1], [A1, A0]):
        rc, node, tet, niw, idx, w = _locate(_poly_tables(order), x)
        assert rc == 0 and tet[0] == 0


def test_fallback_scores_circumsphere_rejections_by_minus_d_squared():
    """A small tetrahedron whose circumsphere excludes the point scores -d^2 = -0.140625, larger
    than the other two's smallest weights (-1, -0.5): it is chosen, then its weights are computed
    without the shortcut."""
    x = [1.75, 1.625, 1.625]
    c, _ = (lambda r: (r[0], r[1]))(__import__("conftest").tet_record(SMALL))
    d2 = np.sum((c[:3] - x) ** 2)
    assert d2 == 0.140625 and d2 > c[3] ** 2
    t = _poly_tables([A0, A1, SMALL])
    rc, node, tet, niw, idx, w = _locate(t, x)
    assert rc == 0 and tet[0] == 2
    assert np.allclose(np.sort(w[0, :niw[0]]), np.sort(_bary(SMALL, x)[np.abs(_bary(SMALL, x)) > 1e-12]))
    # without the small tetrahedron the choice is A1 again
    rc, node, tet, *_ = _locate(_poly_tables([A0, A1]), x)
    assert tet[0] == 1
