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


def bare_grid(cfg, volume_fraction=None, always_triangulate=False):
    """The BZTrellisQdc of a configuration, not filled (for payloads that are
    generated and filled on the device: configuration 3's complex128
    eigenvectors alone would be 83 GB on the host)."""
    c = CONFIGS[cfg]
    bz = _b.BrillouinZone(lattice(cfg))
    vf = c["volume_fraction"] if volume_fraction is None else volume_fraction
    return _b.BZTrellisQdc(bz, vf * bz.ir_polyhedron.volume, always_triangulate)


def layout_tables(g, cfg):
    """Tables of an unfilled grid plus its configuration's payload layout (scalar
    eigenvalues, one Gamma-rotating 3-vector per atom in real-lattice units) and
    GammaTable: what brille_grid_fill_ex needs besides the payload itself."""
    c = CONFIGS[cfg]
    n_atoms = c["modes"] // 3
    t = dict(g.tables())
    t.update(g.gamma_tables())
    t["values_elements"] = [1, 0, 0]
    t["vectors_elements"] = [0, 3 * n_atoms, 0]
    t["values_rotates_like"], t["values_length_unit"] = 0, 3  # the defaults fill gives scalars
    t["vectors_rotates_like"], t["vectors_length_unit"] = 2, 3
    t["branches"] = c["modes"]
    for k in ("values", "vectors"):
        t.pop(k, None)
    return t


def q_points(cfg, n=None, stream=2):
    c = CONFIGS[cfg]
    n = c["n_q"] if n is None else n
    return rng(cfg, stream).uniform(-c["box"], c["box"], size=(int(n), 3))


# ---------------------------------------------------------------------------
# configuration 2: many uneven grids for the batched launch

BATCH_SPACEGROUPS = ("221", "139", "62", "14", "194")
BATCH_MONOCLINIC_BETA = (95.0, 115.0)  # beta is drawn from this range
BATCH_PAYLOAD_CAP = 256 << 20          # bytes of complex128 eigenvectors per grid (caps V for many atoms)


def _batch_lattice_parameters(sg, r):
    a, b, c = r.uniform(3.0, 8.0, 3)
    if sg == "221":
        return (a, a, a), (90.0, 90.0, 90.0)
    if sg == "139":
        return (a, a, c), (90.0, 90.0, 90.0)
    if sg == "62":
        return (a, b, c), (90.0, 90.0, 90.0)
    if sg == "14":
        return (a, b, c), (90.0, float(r.uniform(*BATCH_MONOCLINIC_BETA)), 90.0)
    return (a, a, c), (90.0, 90.0, 120.0)


def _orbit(sym, x):
    pos = []
    for k in range(sym.W.shape[0]):
        p = (sym.W[k] @ x + sym.w[k]) % 1.0
        p[np.isclose(p, 1.0, atol=1e-9)] = 0.0
        if not any(np.allclose(p, q, atol=1e-8) for q in pos):
            pos.append(p)
    return pos


def _batch_basis(sg, target, r):
    """A symmetry-closed basis of at most 24 atoms and about `target` atoms:
    whole orbits of special and general positions."""
    sym = _b.Symmetry(sg)
    cands = [lambda: np.zeros(3), lambda: np.full(3, 0.5), lambda: np.array([0.0, 0.5, 0.5]),
             lambda: np.array([0.5, 0.0, 0.0]), lambda: np.full(3, 0.25),
             lambda: np.array([r.uniform(0.05, 0.45), 0.0, 0.0]),
             lambda: np.full(3, r.uniform(0.05, 0.45)),
             lambda: np.array([r.uniform(0.05, 0.45), r.uniform(0.05, 0.45), 0.0]),
             lambda: np.array([r.uniform(0.05, 0.45), 0.25, r.uniform(0.05, 0.45)]),
             lambda: r.uniform(0.05, 0.95, 3)]
    pos, typ = [], []
    for attempt in range(400):
        if len(pos) >= target:
            break
        orb = _orbit(sym, cands[r.integers(len(cands))]())
        if len(pos) + len(orb) > 24 or any(np.allclose(p, q, atol=1e-6) for p in orb for q in pos):
            continue
        pos.extend(orb)
        typ.extend([len(set(typ))] * len(orb))
    if not pos:
        pos, typ = [np.zeros(3)], [0]
    return np.array(pos), np.array(typ)


def batch_specs(n_grids=64, v_range=(2e3, 2e5), cfg=2, payload_cap=BATCH_PAYLOAD_CAP):
    """Per-grid parameters of configuration 2: space group cycling
    BATCH_SPACEGROUPS, lattice lengths U(3,8) A (monoclinic beta U(95,115)
    deg), a symmetry-closed basis of 1-24 atoms (target log-uniform), 3A modes
    and a vertex count log-uniform in v_range -- optionally capped so that the
    complex128 eigenvectors stay under payload_cap bytes (None: no cap)."""
    r = rng(cfg, 0)
    specs = []
    for gi in range(n_grids):
        sg = BATCH_SPACEGROUPS[gi % len(BATCH_SPACEGROUPS)]
        lengths, angles = _batch_lattice_parameters(sg, r)
        target_atoms = int(round(np.exp(r.uniform(0.0, np.log(24.0)))))
        pos, typ = _batch_basis(sg, target_atoms, r)
        A = len(pos)
        v_target = float(np.exp(r.uniform(np.log(v_range[0]), np.log(v_range[1]))))
        if payload_cap is not None:
            v_target = min(v_target, payload_cap / (3 * A * 3 * A * 16.0))
        v_target = max(v_target, min(v_range[0], 200.0))
        specs.append(dict(index=gi, sg=sg, lengths=lengths, angles=angles, pos=pos, typ=typ, atoms=A,
                          modes=3 * A, v_target=v_target))
    return specs


def batch_grid(spec):
    """The unfilled BZTrellisQdc of one batch grid: the node volume is solved
    from one trial build for the spec's vertex count."""
    lat = _b.Direct(np.array(spec["lengths"]), np.array(spec["angles"]), spec["sg"],
                    _b.Basis(spec["pos"], spec["typ"]))
    bz = _b.BrillouinZone(lat)
    trial_nv = bz.ir_polyhedron.volume / 300.0
    trial = _b.BZTrellisQdc(bz, trial_nv)
    return _b.BZTrellisQdc(bz, trial_nv * trial.rlu.shape[0] / spec["v_target"])


def batch_payload(g, spec, cfg=2):
    """Payload of one batch grid: smooth seeded eigenvalues (V, B, 1) and random
    unit eigenvectors (V, B, A, 3) complex128, each mode's scattered about a
    direction of its own (u_b + 0.3 noise, normalised), so that overlaps
    between vertices stay far from zero and their antiphases are well
    conditioned in FP32 (not symmetrised: the batch exercises scheduling over
    uneven grids, not the physics)."""
    rlu = g.rlu
    V, B, A = rlu.shape[0], spec["modes"], spec["atoms"]
    rv = rng(cfg, 100 + spec["index"])
    Gk = rv.integers(-2, 3, size=(B, 2, 3)).astype(np.float64)
    Ak = rv.uniform(0.5, 5.0, size=(B, 2))
    vals = np.empty((V, B, 1))
    for j in range(B):
        vals[:, j, 0] = np.cos(2 * np.pi * rlu @ Gk[j].T) @ Ak[j]
    u = rv.normal(size=(B, A, 3)) + 1j * rv.normal(size=(B, A, 3))
    u /= np.linalg.norm(u.reshape(B, -1), axis=1)[:, None, None]
    vecs = rv.normal(size=(V, B, A, 3)) + 1j * rv.normal(size=(V, B, A, 3))
    vecs *= 0.3 / np.sqrt(2 * 3 * A)
    vecs += u[None]
    vecs /= np.linalg.norm(vecs.reshape(V, B, -1), axis=2)[:, :, None, None]
    return vals, vecs


def batch_layout_tables(g, spec):
    """Tables of a batch grid plus its payload layout (scalar eigenvalues, one
    Gamma-rotating 3-vector per atom in real-lattice units) and GammaTable:
    with batch_payload, what brille_grid_fill_ex needs."""
    A = spec["atoms"]
    t = dict(g.tables())
    t.update(g.gamma_tables())
    t.update(values_elements=[1, 0, 0], vectors_elements=[0, 3 * A, 0], values_rotates_like=0,
             values_length_unit=3, vectors_rotates_like=2, vectors_length_unit=3, branches=3 * A)
    for k in ("values", "vectors"):
        t.pop(k, None)
    return t


def batch_queries(n_grids=64, n_q=10_000_000, box=3.0, cfg=2):
    """Q of configuration 2: n_q points uniform in [-box, box]^3 split across
    grids with Dirichlet(0.3) weights; grid g owns Q[offsets[g]:offsets[g+1]]."""
    w = rng(cfg, 1).dirichlet(np.full(n_grids, 0.3))
    counts = np.floor(w * n_q).astype(np.int64)
    counts[np.argmax(counts)] += int(n_q) - int(counts.sum())
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    Q = rng(cfg, 2).uniform(-box, box, size=(int(n_q), 3))
    return Q, offsets


def batch_grids(n_grids=64, v_range=(2e3, 2e5), n_q=10_000_000, box=3.0, cfg=2, verbose=None,
                payload_cap=BATCH_PAYLOAD_CAP):
    """Configuration 2 held on the host: n_grids filled BZTrellisQdc of uneven
    vertex and mode counts and their Q (batch_specs, batch_grid,
    batch_payload, batch_queries).  Returns (grids, Q, offsets).  The default
    cap keeps every grid's complex128 payload on the host small; the full
    configuration streams grid by grid instead (payload_cap=None)."""
    grids = []
    for spec in batch_specs(n_grids, v_range, cfg, payload_cap):
        g = batch_grid(spec)
        vals, vecs = batch_payload(g, spec, cfg)
        A = spec["atoms"]
        g.fill(vals, np.array([1, 0, 0]), vecs, np.array([0, 3 * A, 0, 2, 3]))
        grids.append(g)
        if verbose:
            verbose(f"grid {spec['index']}: sg {spec['sg']} A={A} B={3 * A} V={g.rlu.shape[0]} "
                    f"(target {spec['v_target']:.0f})")
    Q, offsets = batch_queries(n_grids, n_q, box, cfg)
    return grids, Q, offsets


def shuffled_modes(tables, seed=0, spread=10.0):
    """A payload whose modes never cross, then shuffled independently at every
    vertex, for the mode sort.

    Mode b has eigenvalue spread * b + a_b . x (x Cartesian, |a_b| < 1 over
    the zone) and eigenvector phase_v,b (u_b + 0.05 g_b . x), with unit u_b
    drawn at random (nearly orthogonal), a random phase per vertex and mode,
    and the element layout of `tables`.  Returns (values, vectors, order):
    the shuffled payload and order[v, a], the original mode stored at vertex
    v as mode a; the sort of an exact assignment maps mode a of vertex i to
    the mode b of vertex j with order[j, b] == order[i, a]."""
    r = np.random.default_rng(np.random.SeedSequence([2026, 99, int(seed)]))
    x = np.asarray(tables["vertices"], dtype=np.float64)
    V = x.shape[0]
    B = int(tables["branches"])
    sv = int(np.sum(tables["values_elements"]))
    sc = int(np.sum(tables["vectors_elements"]))
    scale = max(1e-12, float(np.max(np.abs(x))))
    a = r.normal(size=(B, 3)) * 0.3 / scale
    vals = spread * np.arange(B)[None, :] + x @ a.T
    values = np.repeat(vals[:, :, None], sv, axis=2)
    u = r.normal(size=(B, sc)) + 1j * r.normal(size=(B, sc))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    gk = (r.normal(size=(B, sc, 3)) + 1j * r.normal(size=(B, sc, 3))) * 0.05 / scale
    vec = u[None] + np.einsum("bek,vk->vbe", gk, x)
    vec *= np.exp(2j * np.pi * r.random((V, B)))[:, :, None]
    order = np.argsort(r.random((V, B)), axis=1)
    rows = np.arange(V)[:, None]
    return np.ascontiguousarray(values[rows, order]), np.ascontiguousarray(vec[rows, order]), order
