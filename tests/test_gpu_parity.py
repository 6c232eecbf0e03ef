"""Parity of the HIP path with the oracle, through the C ABI (needs an MI355X).

Indices, tau, q and the interpolation weights are compared bit for bit; FP64
values and eigenvectors within 1e-12 relative; the FP32/complex64 path within
1e-5 relative of the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from conftest import linear_field_grid, nacl_grid, phase_free_vectors
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EXACT = ("ir_q", "tau", "rot", "invrot", "node", "tet", "n_iw", "iw_index", "iw_weight", "status")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def adversarial_points(g, n_random, cfg, stream=3):
    """Random Q plus points on zone faces, edges and vertices, on wedge planes, on knot planes,
    at Gamma and nudged by one ulp either side."""
    bz = g.BrillouinZone
    t = g.tables()
    pts = [S.q_points(cfg, n_random, stream=stream)]
    B = t["outer_to_xyz"].reshape(3, 3)
    Binv = np.linalg.inv(B)
    special = [np.zeros((1, 3)), np.asarray(bz.vertices), np.asarray(bz.points), np.asarray(bz.ir_vertices)]
    ir = np.asarray(bz.ir_vertices)
    # midpoints of irreducible edges and faces
    special.append((ir[:, None, :] + ir[None, :, :]).reshape(-1, 3) / 2)
    # points on knot planes inside the irreducible zone
    rng = np.random.default_rng(stream)
    w = rng.dirichlet(np.ones(len(ir)), size=200)
    inside = w @ ir
    x = inside @ B.T
    for d in range(3):
        k = t["knots"][d]
        snap = x.copy()
        snap[:, d] = k[np.clip(np.searchsorted(k, snap[:, d]), 0, len(k) - 1)]
        special.append(snap @ Binv.T)
    special.append(inside)
    sp = np.vstack(special)
    Rs = t["rotations"].reshape(-1, 3, 3)
    moved = [sp]
    for j in range(0, len(Rs), max(1, len(Rs) // 6)):
        moved.append(sp @ Rs[j] + rng.integers(-2, 3, size=(1, 3)))
    sp = np.vstack(moved)
    nudged = [sp, np.nextafter(sp, np.inf), np.nextafter(sp, -np.inf)]
    return np.vstack(pts + nudged)


def compare(g_tables, Q, precision, kernel, no_move=False):
    dev = capi.Grid(g_tables)
    rc, ov, oc, oaux, bad = O.ir_interpolate_at(g_tables, Q, no_move=no_move, nthreads=8)
    try:
        gv, gc, gaux = dev.ir_interpolate_at(Q, precision=precision, kernel=kernel, no_move=no_move, aux=True)
        grc = 0
    except capi.BrilleError as e:
        grc = e.code
        gv = gc = gaux = None
    return rc, ov, oc, oaux, grc, gv, gc, gaux


CASES = [(0, None), (1, None), (4, 1e-4)]


@pytest.mark.parametrize("cfg,vf", CASES)
@pytest.mark.parametrize("kernel", [capi.KERNEL_PLAIN, capi.KERNEL_TUNED])
def test_fp64_parity(cfg, vf, kernel):
    say(f"fp64 parity cfg={cfg} vf={vf} kernel={kernel}")
    g = S.grid(cfg, volume_fraction=vf)
    t = g.tables()
    Q = adversarial_points(g, 20000, cfg)
    rc, ov, oc, oaux, grc, gv, gc, gaux = compare(t, Q, capi.FP64, kernel)
    assert rc == 0 and grc == 0
    for k in EXACT:
        assert np.array_equal(oaux[k], gaux[k]), k
    scale_v = np.max(np.abs(ov))
    assert np.max(np.abs(gv - ov)) <= 1e-12 * scale_v
    scale_c = max(np.max(np.abs(oc)), 1e-300)
    assert np.max(np.abs(gc - oc)) <= 1e-12 * scale_c


@pytest.mark.parametrize("cfg,vf", CASES)
def test_fp32_path_indices_exact_values_within_1e5(cfg, vf):
    say(f"fp32 parity cfg={cfg} vf={vf}")
    g = S.grid(cfg, volume_fraction=vf)
    t = g.tables()
    Q = adversarial_points(g, 20000, cfg, stream=4)
    rc, ov, oc, oaux, grc, gv, gc, gaux = compare(t, Q, capi.FP32, capi.KERNEL_TUNED)
    assert rc == 0 and grc == 0
    for k in EXACT:
        assert np.array_equal(oaux[k], gaux[k]), k
    # eigenvalues: error relative to each mode's eigenvalue scale (the synthetic cosine series
    # passes through zero, where a pointwise ratio is meaningless)
    scale = np.max(np.abs(ov), axis=0, keepdims=True)
    rel_v = np.max(np.abs(gv - ov) / scale)
    assert rel_v <= 1e-5, rel_v
    # eigenvectors: error relative to the norm of each mode's vector
    nrm = np.sqrt(np.sum(np.abs(oc.reshape(oc.shape[0], oc.shape[1], -1)) ** 2, axis=2))
    err = np.sqrt(np.sum(np.abs((gc - oc).reshape(oc.shape[0], oc.shape[1], -1)) ** 2, axis=2))
    assert np.max(err / np.maximum(nrm, 1e-30)) <= 1e-5


def test_nacl_golden_on_gpu(nacl):
    """wrap/tests/test_5_gamma.py:95-178 through the pybind11 surface (drives the C ABI)."""
    say("nacl golden on gpu")
    d = nacl
    lat, bz, g, perm, q_nu, _ = nacl_grid(d)
    for prec, tol in (("fp64", 1e-8), ("fp32", 1e-4)):
        v, c = g.ir_interpolate_at(q_nu, precision=prec)
        v = np.squeeze(v)
        assert np.allclose(v, d["euphonic_values"], rtol=tol, atol=tol * 10)
        br = np.einsum("ba,ijkb->ijka", d["basis_vectors"], c.astype(np.complex128))
        assert np.allclose(phase_free_vectors(br, d["euphonic_vectors"]), d["euphonic_vectors"], atol=tol * 10)


def test_linear_field_on_gpu():
    """src/tests/trellis_test.cpp:76-137 on the GPU, plus rip_recip for arbitrary Q."""
    say("linear field on gpu")
    lat, bz, g = linear_field_grid(3e-4)
    Q = np.random.default_rng(6).uniform(-3, 3, size=(5000, 3))
    dev = capi.Grid(g.tables())
    v, c, aux = dev.ir_interpolate_at(Q, aux=True)
    B = np.asarray(lat.reciprocal_vectors)
    assert np.max(np.abs(v[:, 0, :] - (Q - aux["tau"]) @ B.T)) < 2e-10


def test_errors_match_the_reference_messages(nacl):
    say("error paths")
    g = S.grid(0)
    dev = capi.Grid(g.tables())
    with pytest.raises(capi.BrilleError, match="Not all points inside Brillouin zone") as e:
        dev.ir_interpolate_at(np.array([[0.1, 0.2, 0.3], [np.nan, 0.0, 0.0]]))
    assert e.value.status == "OUTSIDE_BZ"
    v, c = dev.ir_interpolate_at(np.zeros((0, 3)))
    assert v.shape[0] == 0
    # an unfilled handle reports before launching
    bare = {k: val for k, val in g.tables().items() if k not in ("values", "vectors")}
    h = capi.Grid(bare)
    with pytest.raises(capi.BrilleError, match="must be filled"):
        h.ir_interpolate_at(np.zeros((3, 3)))
    # unsupported RotatesLike / LengthUnit combinations (wrap/tests/test_5_gamma.py:223-240)
    lat, bz, gn, perm, q_nu, _ = nacl_grid(nacl)
    for combo in [(0, 1), (1, 0), (2, 2), (2, 4)]:
        gn.fill(nacl["grid_values"][perm], nacl["grid_values_elements"].astype(int), nacl["grid_values_weights"],
                nacl["grid_vectors"][perm], np.array([0, 24, 0, *combo, 0, 0]), nacl["grid_vectors_weights"], False)
        with pytest.raises(RuntimeError):
            gn.ir_interpolate_at(nacl["q_nu"])


def test_no_move_and_device_memory_path():
    say("no_move and device memory")
    g = S.grid(1)
    t = g.tables()
    dev = capi.Grid(t)
    Q = S.q_points(1, 4096, stream=5)
    v, c, aux = dev.ir_interpolate_at(Q, aux=True)
    # the folded points interpolated without moving give the unrotated sums
    v2, c2 = dev.ir_interpolate_at(aux["ir_q"], no_move=True)
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, aux["ir_q"], no_move=True)
    assert np.max(np.abs(v2 - ov)) <= 1e-12 * np.max(np.abs(ov))
    # device-resident inputs and outputs give the same bits as the staged host call
    dQ = capi.DeviceArray.from_host(Q)
    vs, cs = dev.out_shapes(len(Q))
    dv = capi.DeviceArray(vs, np.float64)
    dc = capi.DeviceArray(cs, np.complex128)
    import ctypes as C
    o = capi.RunOpts(0, capi.FP64, 0, capi.MEM_DEVICE, capi.KERNEL_AUTO, 1)
    capi.check(capi.lib().brille_ir_interpolate_at(dev.h, dQ.ptr, len(Q), C.byref(o), dv.ptr, dc.ptr, None, None))
    assert np.array_equal(dv.to_host(), v)
    assert np.array_equal(dc.to_host(), c)


def test_ir_moveinto_entry_point():
    say("ir_moveinto entry point")
    g = S.grid(1)
    t = g.tables()
    dev = capi.Grid(t)
    Q = S.q_points(1, 10000, stream=6)
    rc, q, tau, r, ir, st = dev.ir_moveinto(Q)
    orc, oq, otau, orr, oir, ost = O.ir_moveinto(t, Q)
    assert rc == orc == 0
    for a, b_ in ((q, oq), (tau, otau), (r, orr), (ir, oir), (st, ost)):
        assert np.array_equal(a, b_)


@pytest.mark.parametrize("kernel", [capi.KERNEL_PLAIN, capi.KERNEL_TILE, capi.KERNEL_BINNED])
def test_first_zone_interpolate_at(kernel):
    """brille_interpolate_at (BrillouinZoneTrellis3::interpolate_at: moveinto only,
    no rotation) against the oracle's first-zone mode."""
    say(f"interpolate_at kernel={kernel}")
    g = S.grid(1)
    t = g.tables()
    dev = capi.Grid(t)
    bz = g.BrillouinZone
    ir = np.asarray(bz.ir_vertices)
    r = np.random.default_rng(8)
    q = r.dirichlet(np.ones(len(ir)), size=5000) @ ir
    Q = q + r.integers(-2, 3, size=(5000, 3))
    aux = dict(tau=False) if kernel == capi.KERNEL_BINNED else True
    gv, gc, ga = dev.interpolate_at(Q, precision=capi.FP64, kernel=kernel, aux=aux)
    rc, ov, oc, oa, _ = O.ir_interpolate_at(t, Q, nthreads=8, first_zone=True)
    assert rc == 0
    for k in ga:
        np.testing.assert_array_equal(ga[k], oa[k], err_msg=k)
    np.testing.assert_allclose(gv, ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max())
    np.testing.assert_allclose(gc, oc, rtol=1e-12, atol=1e-12)
    # first-zone images outside the wedge fail as in the reference, at the same first point
    Qb = S.q_points(1, 3000, stream=42)
    rc, *_ , bad = O.ir_interpolate_at(t, Qb, nthreads=8, first_zone=True)
    with pytest.raises(capi.BrilleError) as e:
        dev.interpolate_at(Qb, precision=capi.FP32, kernel=kernel)
    assert e.value.status == "NULL_NODE"
    assert f"{Qb[bad][0]:.17g}"[:12] in str(e.value) or str(bad) in str(e.value)


@pytest.mark.parametrize("kernel", [capi.KERNEL_PLAIN, capi.KERNEL_TILE, capi.KERNEL_BINNED])
def test_earlier_stage_errors_win(kernel):
    """A fold-in failure (NaN: outside the zone) at a later index is the call's error even when a
    null node comes first, as the reference folds every point before locating any
    (src/bz_trellis.hpp:164-168)."""
    say(f"error precedence kernel={kernel}")
    g = S.grid(1)
    t = g.tables()
    dev = capi.Grid(t)
    Qb = S.q_points(1, 3000, stream=42)
    rc, *_, bad = O.ir_interpolate_at(t, Qb, first_zone=True)
    Q = np.vstack([Qb[bad:bad + 1], Qb[:50], [[np.nan, 0.0, 0.0]], Qb[:5]])
    orc, *_, obad = O.ir_interpolate_at(t, Q, first_zone=True)
    assert orc == 3 and obad == 51
    with pytest.raises(capi.BrilleError) as e:
        dev.interpolate_at(Q, precision=capi.FP32, kernel=kernel)
    assert e.value.status == "OUTSIDE_BZ" and "nan" in str(e.value)
