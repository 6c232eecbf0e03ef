"""BZMeshQdc / BZNestQdc on the GPU against the oracle (needs an MI355X).

The mesh grids run the same device path as BZTrellisQdc with every node
triangulated: TILE and BINNED against the oracle on the same tables, indices
bit for bit, values and vectors within 1e-12, for points anywhere in
reciprocal space (fold-in, location in a tetrahedron, rotation)."""
import numpy as np
import pytest

import brille_amd as b
from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["mesh", "nest"])
def test_mesh_grid_matches_the_oracle(kind):
    g0 = S.grid(1)  # config 1's lattice and payload generator
    bz = g0.BrillouinZone
    g = b.BZMeshQdc(bz, max_size=0.001) if kind == "mesh" else b.BZNestQdc(bz, 2000)
    x = np.asarray(g.rlu)
    r = S.rng(1, 70)
    W = np.asarray(bz.pointgroup.W)
    vals = S.symmetric_values(x, W, 6, r)
    vecs = S.hermitian_vectors(x, 2, r)
    g.fill(vals, np.array([1, 0, 0]), vecs.reshape(x.shape[0], 6, 6), np.array([0, 6, 0, 2, 3]))
    t = g.tables()
    dev = capi.Grid(t)
    Q = S.q_points(1, 50_000, stream=71)
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=16)
    assert rc == 0
    for kernel in (capi.KERNEL_TILE, capi.KERNEL_BINNED):
        aux = dict(tau=kernel != capi.KERNEL_BINNED)
        gv, gc, gaux = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=kernel, aux=aux)
        for k in gaux:
            np.testing.assert_array_equal(gaux[k], oaux[k], err_msg=k)
        assert (gaux["tet"] != 0xFFFFFFFF).all()
        assert np.max(np.abs(gv - ov)) <= 1e-12 * np.abs(ov).max()
        assert np.max(np.abs(gc - oc)) <= 1e-12 * np.abs(oc).max()
    # the Python surface gives the same
    pv, pc = g.ir_interpolate_at(Q)
    np.testing.assert_array_equal(pv, gv)
