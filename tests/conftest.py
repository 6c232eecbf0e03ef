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
