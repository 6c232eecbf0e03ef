"""Synthetic inputs of the benchmark configurations (no datasets are needed).

Every configuration is generated from ``numpy.random.default_rng(SeedSequence([2026, cfg, stream]))``.

* eigenvalues: omega_j(q) = (1/|G|) sum_R sum_k A_jk cos(2 pi G_k . R^T q) with A_jk ~ U(0.5, 5)
  and G_k drawn from {-2..2}^3, so the data carry the point-group symmetry of the lattice;
* eigenvectors: eigenvectors of the Hermitian matrix H0 + sum_k cos(2 pi G_k . q) H_k with
  seed-drawn Hermitian H's (3A x 3A for A atoms);
* Q: uniform in a cube of reciprocal lattice units.
"""
import numpy as np

from . import _brille as _b

# lattice, symmetry and atoms of each configuration
CONFIGS = {
    0: dict(name="Fd-3m cubic a=4.96", lengths=(4.96, 4.96, 4.96), angles=(90, 90, 90), spacegroup="227",
            atoms=None, modes=6, volume_fraction=1e-2, n_q=100_000, box=3.0, vectors=False),
    1: dict(name="P6_3/mmc hexagonal a=3.2 c=5.2, 2 atoms", lengths=(3.2, 3.2, 5.2), angles=(90, 90, 120),
            spacegroup="194", atoms=([[1 / 3, 2 / 3, 0.25], [2 / 3, 1 / 3, 0.75]], [0, 0]), modes=6,
            volume_fraction=1e-2, n_q=1_000_000, box=4.0, vectors=True),
    3: dict(name="P2_1/c monoclinic a=5.1 b=7.3 c=9.0 beta=104, 24 atoms", lengths=(5.1, 7.3, 9.0),
            angles=(90, 104, 90), spacegroup="14", atoms="general24", modes=72, volume_fraction=1.9e-7 / 0.1907,
            n_q=100_000_000, box=5.0, vectors=True),
    4: dict(name="P6_3/mmc hexagonal a=3.2 c=5.2, 2 atoms, fine grid", lengths=(3.2, 3.2, 5.2),
            angles=(90, 90, 120), spacegroup="194", atoms=([[1 / 3, 2 / 3, 0.25], [2 / 3, 1 / 3, 0.75]], [0, 0]),
            modes=6, volume_fraction=1e-5, n_q=100_000_000, box=4.0, vectors=True),
}


def rng(cfg, stream):
    return np.random.default_rng(np.random.SeedSequence([2026, int(cfg), int(stream)]))


def _general_positions(sg, n_orbits, r):
    """Atoms of n_orbits general-position orbits of a space group."""
    W = sg.W
    w = sg.w
    pos, typ = [], []
    for o in range(n_orbits):
        x = r.uniform(0.05, 0.95, 3)
        for k in range(W.shape[0]):
            p = (W[k] @ x + w[k]) % 1.0
            if not any(np.allclose(p, q, atol=1e-8) for q in pos):
                pos.append(p)
                typ.append(o)
    return np.array(pos), np.array(typ)


def lattice(cfg):
    c = CONFIGS[cfg]
    atoms = c["atoms"]
    if atoms is None:
        return _b.Direct(np.array(c["lengths"]), np.array(c["angles"]), c["spacegroup"])
    if atoms == "general24":
        sg = _b.Symmetry(c["spacegroup"])
        pos, typ = _general_positions(sg, 6, rng(cfg, 9))
        basis = _b.Basis(pos, typ)
    else:
        basis = _b.Basis(np.array(atoms[0]), np.array(atoms[1]))
    return _b.Direct(np.array(c["lengths"]), np.array(c["angles"]), c["spacegroup"], basis)


def symmetric_values(rlu, W, modes, r, n_g=8):
    """(V, modes, 1) eigenvalues symmetrised over the point group W (|G| x 3 x 3)."""
    A = r.uniform(0.5, 5.0, size=(modes, n_g))
    G = r.integers(-2, 3, size=(modes, n_g, 3)).astype(np.float64)
    out = np.zeros((rlu.shape[0], modes))
    for R in W:
        q = rlu @ R  # rows of R^T q
        for j in range(modes):
            out[:, j] += np.cos(2 * np.pi * q @ G[j].T) @ A[j]
    out /= W.shape[0]
    out.sort(axis=1)
    return out[:, :, None]


def hermitian_vectors(rlu, n_atoms, r, n_g=4):
    """(V, 3A, A, 3) complex eigenvectors of a q-dependent Hermitian matrix."""
    m = 3 * n_atoms
    H = r.normal(size=(n_g + 1, m, m)) + 1j * r.normal(size=(n_g + 1, m, m))
    H = H + np.conj(np.swapaxes(H, 1, 2))
    G = r.integers(-2, 3, size=(n_g, 3)).astype(np.float64)
    out = np.empty((rlu.shape[0], m, n_atoms, 3), np.complex128)
    step = 65536
    for s in range(0, rlu.shape[0], step):
        q = rlu[s:s + step]
        D = np.broadcast_to(H[0], (q.shape[0], m, m)).copy()
        c = np.cos(2 * np.pi * q @ G.T)
        D += np.einsum("vk,kab->vab", c, H[1:])
        _, vec = np.linalg.eigh(D)
        out[s:s + step] = np.swapaxes(vec, 1, 2).reshape(q.shape[0], m, n_atoms, 3)
    return out


def grid(cfg, volume_fraction=None, always_triangulate=False, with_vectors=None):
    """BZTrellisQdc of a configuration, filled with its synthetic data."""
    c = CONFIGS[cfg]
    lat = lattice(cfg)
    bz = _b.BrillouinZone(lat)
    vf = c["volume_fraction"] if volume_fraction is None else volume_fraction
    g = _b.BZTrellisQdc(bz, vf * bz.ir_polyhedron.volume, always_triangulate)
    r = rng(cfg, 0)
    rlu = g.rlu
    vals = symmetric_values(rlu, lat.pointgroup.W, c["modes"], r)
    want_vec = c["vectors"] if with_vectors is None else with_vectors
    if want_vec:
        n_atoms = c["modes"] // 3
        vecs = hermitian_vectors(rlu, n_atoms, rng(cfg, 1))
        g.fill(vals, np.array([1, 0, 0]), vecs, np.array([0, 3 * n_atoms, 0, 2, 3]))
    else:
        vecs = np.zeros((rlu.shape[0], c["modes"], 1), np.complex128)
        g.fill(vals, np.array([1, 0, 0]), vecs, np.array([1, 0, 0]))
    return g


def q_points(cfg, n=None, stream=2):
    c = CONFIGS[cfg]
    n = c["n_q"] if n is None else n
    return rng(cfg, stream).uniform(-c["box"], c["box"], size=(int(n), 3))
