import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def nacl():
    """The reference's NaCl fixture (wrap/tests/test_5_gamma.npz), committed as data."""
    return dict(np.load(os.path.join(GOLDEN, "nacl_gamma.npz")))


def nacl_grid(d):
    """Lattice, zone and grid of the NaCl fixture, filled as the reference's test does
    (wrap/tests/test_5_gamma.py:71-129)."""
    import brille_amd as b
    lat = b.Lattice(d["basis_vectors"], b.Symmetry(d["spacegroup_mat"], d["spacegroup_vec"]),
                    b.Basis(d["atom_positions"], d["atom_index"]))
    bz = b.BrillouinZone(lat)
    g = b.BZTrellisQdc(bz, float(d["grid_max_volume"]), bool(d["grid_always_triangulate"]))
    perm = np.hstack([np.argwhere(np.all(np.isclose(d["grid_rlu"], x), axis=1)) for x in g.rlu]).flatten()
    g.fill(d["grid_values"][perm], d["grid_values_elements"].astype(int), d["grid_values_weights"],
           d["grid_vectors"][perm], np.array([0, 24, 0, 2, 3, 0, 0]), d["grid_vectors_weights"], False)
    q_ir = d["grid_rlu"][4:5]
    q_nu = np.einsum("xji,aj->xi", lat.pointgroup.W, q_ir)
    p2 = np.array([np.squeeze(np.argwhere(np.all(np.isclose(q_nu, x), axis=1))) for x in d["q_nu"]])
    return lat, bz, g, perm, q_nu[p2], p2


def linear_field_grid(max_volume=1e-4):
    """Nb (Im-3m, a=3.2598) grid filled with each vertex's own xyz as a reciprocal-lattice
    vector (src/tests/trellis_test.cpp:76-137)."""
    import brille_amd as b
    lat = b.Direct(np.array([3.2598] * 3), np.array([90.0] * 3), "-I 4 2 3")
    bz = b.BrillouinZone(lat)
    g = b.BZTrellisQdc(bz, max_volume)
    xyz = np.asarray(g.invA)
    g.fill(xyz[:, None, :], np.array([0, 3, 0, 0, 4]), np.zeros((xyz.shape[0], 1, 1), complex), np.array([1, 0, 0]))
    return lat, bz, g


def phase_free_vectors(a, ref):
    """Remove a per-(point, mode) global phase of `a` against `ref` (test_5_gamma.py:162-175)."""
    ph = np.exp(-1j * np.angle(np.einsum("qmij,qmij->qm", np.conj(ref), a)))
    return np.einsum("ab,abij->abij", ph, a)


def field_grid(spacegroup, lengths, angles, volume_fraction=1e-3):
    """A BZTrellisQdc over the irreducible zone of a lattice with no basis (for closed-form fields)."""
    import brille_amd as b
    lat = b.Direct(np.array(lengths, float), np.array(angles, float), spacegroup)
    bz = b.BrillouinZone(lat)
    g = b.BZTrellisQdc(bz, volume_fraction * bz.ir_polyhedron.volume)
    return lat, bz, g


def fill_field(g, values, elements, vectors=None, vector_elements=(1, 0, 0)):
    """Fill g with (V, k) real values of the given element layout (and complex vectors, zeros by default)."""
    V = values.shape[0]
    if vectors is None:
        vectors = np.zeros((V, 1, 1), complex)
    g.fill(values[:, None, :], np.array(elements), vectors, np.array(vector_elements))
    return g.tables()


def direct_components(lat, cart):
    """Components in the direct (real-lattice) basis of Cartesian rows (n, 3)."""
    A = np.asarray(lat.real_vectors)  # basis vectors as columns
    return np.linalg.solve(A, cart.T).T


def direct_operator(lat, cart):
    """Direct-basis matrices A^-1 M A of Cartesian linear operators (n, 3, 3), flattened to (n, 9)."""
    A = np.asarray(lat.real_vectors)
    return np.einsum("ij,njk,kl->nil", np.linalg.inv(A), cart, A).reshape(-1, 9)


def cartesian(lat, rlu):
    return rlu @ np.asarray(lat.reciprocal_vectors).T


# 90 degree rotation about Cartesian z: J x is an axial-vector field for point groups that fix z
J_Z = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])


# ---------------------------------------------------------------------------
# hand-built trellises: unit knots 0..4 on each axis (4 x 4 x 4 nodes, x fastest),
# a vertex at every knot crossing, Cartesian = rlu (identity basis), so points on
# knot planes are exact

NULL, CUBE, POLY = 2, 3, 4
KNOTS = np.arange(5, dtype=float)


def vertex_id(i, j, k):
    return int(i + 5 * j + 25 * k)


def node_id(i, j, k):
    return int(i + 4 * j + 16 * k)


def cube_corners(i, j, k):
    """Vertex ids of cube (i, j, k) in CubeNode order (000),(100),(110),(010),(101),(001),(011),(111)."""
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (1, 0, 1), (0, 0, 1), (0, 1, 1), (1, 1, 1)]
    return [vertex_id(i + a, j + b, k + c) for a, b, c in off]


def tet_record(xyz):
    """Circumsphere (centre, radius) and volume of a tetrahedron given by four vertices (4, 3)."""
    p = np.asarray(xyz, float)
    A = 2 * (p[1:] - p[0])
    rhs = np.sum(p[1:] ** 2 - p[0] ** 2, axis=1)
    c = np.linalg.solve(A, rhs)
    r = np.linalg.norm(p[0] - c)
    vol = abs(np.linalg.det(p[1:] - p[0])) / 6
    return np.array([*c, r]), vol


def hand_trellis(node_types, polys=None):
    """Tables of a hand-built trellis: node_types maps (i, j, k) -> CUBE or POLY (others NULL);
    polys maps (i, j, k) -> list of tetrahedra, each four (x, y, z) vertices on the vertex lattice.
    The zone part comes from configuration 0 (unused when points are not moved); the payload is
    the linear field (x, y, z) as three scalars per vertex."""
    from brille_amd import synthetic as S
    base = S.grid(0).tables()
    t = {k: v for k, v in base.items()}
    t["outer_to_xyz"] = np.eye(3).ravel()
    t["knots"] = [KNOTS.copy(), KNOTS.copy(), KNOTS.copy()]
    verts = np.array([[i, j, k] for k in range(5) for j in range(5) for i in range(5)], float)
    node_type = np.full(64, NULL, np.uint8)
    node_index = np.zeros(64, np.uint32)
    cubes, tet_v, tet_c, tet_vol, offs = [], [], [], [], [0]
    polys = polys or {}
    for (i, j, k), typ in sorted(node_types.items(), key=lambda kv: node_id(*kv[0])):
        n = node_id(i, j, k)
        node_type[n] = typ
        if typ == CUBE:
            node_index[n] = len(cubes)
            cubes.append(cube_corners(i, j, k))
        else:
            node_index[n] = len(offs) - 1
            for tet in polys[(i, j, k)]:
                tet = np.asarray(tet, float)
                if np.linalg.det(tet[:3] - tet[3]) < 0:  # orient3d(v0..v3) = +6V, as the builder's
                    tet = tet[[0, 1, 3, 2]]
                tet_v.append([vertex_id(*map(int, np.round(p))) for p in tet])
                c, vol = tet_record(tet)
                tet_c.append(c)
                tet_vol.append(vol)
            offs.append(len(tet_v))
    t["node_type"] = node_type
    t["node_index"] = node_index
    t["cube_vertices"] = np.array(cubes, np.uint32).reshape(-1, 8)
    t["poly_tet_offset"] = np.array(offs, np.uint32)
    t["tet_vertices"] = np.array(tet_v, np.uint32).reshape(-1, 4)
    t["tet_circumsphere"] = np.array(tet_c, float).reshape(-1, 4)
    t["tet_volume"] = np.array(tet_vol, float)
    t["vertices"] = verts
    t["values"] = verts[:, None, :].copy()
    t["vectors"] = np.zeros((len(verts), 1, 1), complex)
    t["values_elements"] = [3, 0, 0]
    t["vectors_elements"] = [1, 0, 0]
    t["branches"] = 1
    t.pop("poly_node", None)
    return t
