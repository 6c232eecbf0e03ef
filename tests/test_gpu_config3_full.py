"""Configuration 3 at its stated size (needs an MI355X).

Monoclinic P2_1/c, 24 atoms, 72 modes, about 1.07e6 trellis vertices: the
FP32 payload (72 x 72 complex64 eigenvectors per vertex) is 42.6 GB, and the
complex128 one (83 GB) is not formed at all: the payload is generated as
complex64 and handed to brille_grid_fill_ex as FP32 data, keeping the FP32
rows only.  At this size the wide-row
antiphase table (8 per cube, 4 per tetrahedron, per mode) exceeds the 4 GiB
the sums address, so every antiphase is computed on the fly.

* indices: 1e5 Q uniform in [-5, 5]^3 through TILE (every aux output) and
  BINNED (all but tau), bit for bit against the oracle's table-only fold-in
  and location (BrillouinZone::ir_moveinto, PolyTrellis node and weights);
* values: BINNED == TILE bit for bit on the same points;
* a linear-field payload (values a.x + c; vectors (1 + g.x) u, every Hermitian
  overlap between vertices positive) is reproduced within 1e-5 of its scale
  at points inside the irreducible wedge (not moved, identity operation);
* a call at FP64 is refused: those rows were not kept.
"""
import sys

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LOC = ("rot", "invrot", "node", "tet", "n_iw", "iw_index", "iw_weight", "status")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg3():
    g = S.bare_grid(3)
    t = S.layout_tables(g, 3)
    V = t["vertices"].shape[0]
    B, A = 72, 24
    say(f"config 3 grid: {V} vertices, {t['cube_vertices'].shape[0]} cubes, {t['tet_volume'].size} tetrahedra")
    r = np.random.default_rng(np.random.SeedSequence([2026, 3, 50]))
    a = r.normal(size=(B, 3))
    c = r.uniform(5.0, 10.0, size=B)
    gk = r.normal(size=(B, 3 * A, 3)) * 0.05
    u = r.normal(size=(B, 3 * A)) + 1j * r.normal(size=(B, 3 * A))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = np.asarray(t["vertices"])
    vals = (x @ a.T + c).astype(np.float32)
    vecs = np.empty((V, B, 3 * A), np.complex64)
    gk2 = gk.reshape(-1, 3).T.astype(np.float32)
    u32 = u.astype(np.complex64)
    step = 32768
    for s0 in range(0, V, step):
        amp = 1.0 + (x[s0:s0 + step].astype(np.float32) @ gk2).reshape(-1, B, 3 * A)
        np.multiply(amp, u32, out=vecs[s0:s0 + step])
    say(f"payload generated: {vecs.nbytes / 1e9:.1f} GB of complex64 on the host")
    dev = capi.Grid(t)
    dev.fill_ex(t, vals.ctypes.data, vecs.ctypes.data, capi.MEM_HOST, capi.FP32, capi.KEEP_FP32)
    del vals, vecs
    say("filled (FP32 rows only)")
    fields = dict(a=a, c=c, gk=gk, u=u)
    return g, t, dev, fields


def test_fp64_calls_are_refused(cfg3):
    g, t, dev, f = cfg3
    with pytest.raises(capi.BrilleError, match="FP64"):
        dev.ir_interpolate_at(np.zeros((4, 3)), precision=capi.FP64)


def test_indices_match_the_oracle_and_binned_equals_tile(cfg3):
    g, t, dev, f = cfg3
    Q = S.q_points(3, 100_000, stream=51)
    say("config 3 full: oracle fold-in and location on 1e5 Q")
    rc, q, tau, rot, invrot, st = O.ir_moveinto(t, Q)
    assert rc == 0
    rc, node, tet, niw, idx, w = O.locate(t, O.get_xyz(t, q))
    assert rc == 0
    want = dict(rot=rot, invrot=invrot, node=node, tet=tet, n_iw=niw, iw_index=idx, iw_weight=w, status=st)
    say("config 3 full: TILE FP32")
    rc, tv, tc, ta = dev.run_device(Q, precision=capi.FP32, kernel=capi.KERNEL_TILE,
                                    aux=("ir_q", "tau") + LOC)
    assert rc == 0
    np.testing.assert_array_equal(ta["ir_q"], q)
    np.testing.assert_array_equal(ta["tau"], tau)
    for k in LOC:
        np.testing.assert_array_equal(ta[k], want[k], err_msg=k)
    say("config 3 full: BINNED FP32")
    rc, bv, bc, ba = dev.run_device(Q, precision=capi.FP32, kernel=capi.KERNEL_BINNED, aux=("ir_q",) + LOC)
    assert rc == 0
    for k in LOC:
        np.testing.assert_array_equal(ba[k], want[k], err_msg=k)
    assert np.array_equal(bv, tv)
    assert np.array_equal(bc, tc)
    assert np.isfinite(tc).all() and np.abs(tc).max() > 0.1


def test_linear_field_is_reproduced(cfg3):
    g, t, dev, f = cfg3
    bz = g.BrillouinZone
    ir = np.asarray(bz.ir_vertices)
    r = np.random.default_rng(52)
    q = r.dirichlet(np.ones(len(ir)), size=50_000) @ ir
    say("config 3 full: linear field at points inside the wedge")
    for kernel in (capi.KERNEL_TILE, capi.KERNEL_BINNED):
        rc, v, c, aux = dev.run_device(q, precision=capi.FP32, kernel=kernel, no_move=True, aux=("status",))
        assert rc == 0 and not aux["status"].any()
        x = q @ np.asarray(t["outer_to_xyz"]).reshape(3, 3).T
        want_v = x @ f["a"].T + f["c"]
        amp = 1.0 + np.einsum("nk,bek->nbe", x, f["gk"])
        want_c = amp * f["u"][None]
        ev = np.max(np.abs(v[:, :, 0] - want_v)) / np.abs(want_v).max()
        ec = np.max(np.abs(c.reshape(want_c.shape) - want_c)) / np.abs(want_c).max()
        say(f"  kernel {kernel}: eigenvalue error {ev:.2e}, eigenvector error {ec:.2e} of scale")
        assert ev <= 1e-5 and ec <= 1e-5
