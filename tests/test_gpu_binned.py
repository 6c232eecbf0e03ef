"""The binned pipeline (points ordered by trellis node range before the sums)
against the fused tile kernel and the oracle, through the C ABI.

Both paths compute every point from the same FP64 fold-in record with the same
sums code, so their outputs must agree bit for bit; the FP64 binned outputs
also match the oracle within 1e-12 relative.  Needs an MI355X.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_device(dev, Q, precision, kernel, check=1):
    """Device-memory call; returns (rc, values, vectors)."""
    lib = capi.lib()
    n = Q.shape[0]
    vs, cs = dev.out_shapes(n)
    vd = np.float32 if precision == capi.FP32 else np.float64
    cd = np.complex64 if precision == capi.FP32 else np.complex128
    dQ = capi.DeviceArray.from_host(Q)
    dv = capi.DeviceArray(vs, vd)
    dc = capi.DeviceArray(cs, cd)
    o = capi.RunOpts(0, precision, 0, capi.MEM_DEVICE, kernel, check)
    rc = lib.brille_ir_interpolate_at(dev.h, dQ.ptr, n, C.byref(o), dv.ptr, dc.ptr, None, None)
    return rc, dv.to_host(), dc.to_host()


@pytest.mark.parametrize("cfg,vf,n", [(4, 1e-4, 300_000), (1, None, 300_000), (0, None, 300_000)])
@pytest.mark.parametrize("precision", [capi.FP32, capi.FP64])
def test_binned_equals_tile(cfg, vf, n, precision):
    say(f"binned vs tile cfg{cfg} n={n} precision={precision}")
    g = S.grid(cfg, volume_fraction=vf)
    dev = capi.Grid(g.tables())
    Q = S.q_points(cfg, n, stream=11)
    rb, vb, cb = run_device(dev, Q, precision, capi.KERNEL_BINNED)
    rt, vt, ct = run_device(dev, Q, precision, capi.KERNEL_TILE)
    assert rb == 0 and rt == 0
    np.testing.assert_array_equal(vb, vt)
    np.testing.assert_array_equal(cb, ct)


def test_binned_fp64_matches_oracle():
    say("binned fp64 vs oracle cfg4 vf=1e-4")
    g = S.grid(4, volume_fraction=1e-4)
    t = g.tables()
    dev = capi.Grid(t)
    Q = S.q_points(4, 300_000, stream=12)
    rb, vb, cb = run_device(dev, Q, capi.FP64, capi.KERNEL_BINNED)
    assert rb == 0
    m = 40_000
    rc, ov, oc, _, _ = O.ir_interpolate_at(t, Q[:m], nthreads=8)
    assert rc == 0
    np.testing.assert_allclose(vb[:m], ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max())
    np.testing.assert_allclose(cb[:m], oc, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("payload", ["config4", "wide"])
def test_binned_errors_match_tile(payload):
    """The first failing point (lowest index) and the message are the same on
    both paths; for wide rows the binned sums split each tile's modes over
    several blocks, of which only the first reports."""
    say(f"binned error reporting, {payload}")
    if payload == "config4":
        g = S.grid(4, volume_fraction=1e-4)
        dev = capi.Grid(g.tables())
        Q = S.q_points(4, 300_000, stream=13)
        bad = (123_457, 200_001)
    else:
        grids, Q, _ = S.batch_grids(n_grids=6, v_range=(800, 2000), n_q=20_000)
        sc = [int(sum(g.tables()["vectors_elements"])) for g in grids]
        dev = capi.Grid(grids[int(np.argmax(sc))].tables())
        bad = (12_345, 17_001)
    Q[bad[0]] = [np.nan, 0.0, 0.0]
    Q[bad[1]] = [np.inf, 1.0, 0.0]
    lib = capi.lib()
    msgs = []
    for kernel in (capi.KERNEL_BINNED, capi.KERNEL_TILE):
        rc, v, c = run_device(dev, Q, capi.FP32, kernel)
        buf = C.create_string_buffer(512)
        lib.brille_last_error(buf, 512)
        msgs.append((rc, buf.value.decode()))
        # failed points give zero rows; the others are filled
        assert np.all(v[bad[0]] == 0) and np.all(c[bad[1]] == 0)
        assert np.all(np.abs(v[:1000]).sum(axis=(1, 2)) > 0)
    assert msgs[0] == msgs[1]
    assert msgs[0][0] != 0
    assert "nan" in msgs[0][1].lower()


def test_binned_several_passes():
    """More points than one pass holds (2^24): three passes, the third reusing
    the work space half of the first while the second is summed; the passes
    tile the input exactly."""
    n = (2 << 24) + 4099
    say(f"binned several passes n={n}")
    g = S.grid(0)
    dev = capi.Grid(g.tables())
    Q = S.q_points(0, n, stream=14)
    bad = (2 << 24) + 17
    Q[bad] = [np.nan, 0.0, 0.0]
    rb, vb, cb = run_device(dev, Q, capi.FP32, capi.KERNEL_BINNED)
    rt, vt, ct = run_device(dev, Q, capi.FP32, capi.KERNEL_TILE)
    assert rb == rt != 0
    assert np.all(vb[bad] == 0) and np.all(cb[bad] == 0)
    np.testing.assert_array_equal(vb, vt)
    np.testing.assert_array_equal(cb, ct)
    # the same call twice in a row on one stream gives the same outputs
    rb2, vb2, cb2 = run_device(dev, Q, capi.FP32, capi.KERNEL_BINNED)
    assert rb2 == rb
    np.testing.assert_array_equal(vb2, vb)
    np.testing.assert_array_equal(cb2, cb)


@pytest.mark.parametrize("which", ["fewest", "most"])
def test_wide_rows_binned_equals_tile_and_oracle(which):
    """Many-atom payloads (rows of more than 12 complex elements: one wave per
    (point, mode), antiphases from the wide factor table in the binned path):
    the wide grid of fewest atoms, and the one of most (24 atoms: the binned
    sums take its tiles' Gamma phases in several chunks of points)."""
    grids, Q, offsets = S.batch_grids(n_grids=6, v_range=(800, 2000), n_q=20_000)
    sc = [int(sum(g.tables()["vectors_elements"])) for g in grids]
    if which == "fewest":
        gi = int(np.argmin([c if c > 12 else 1 << 30 for c in sc]))
    else:
        gi = int(np.argmax(sc))
    assert sc[gi] > 12, sc
    t = grids[gi].tables()
    say(f"wide rows: {sc[gi]} complex elements per mode, {int(t['branches'])} modes")
    dev = capi.Grid(t)
    for precision in (capi.FP32, capi.FP64):
        rb, vb, cb = run_device(dev, Q, precision, capi.KERNEL_BINNED)
        rt, vt, ct = run_device(dev, Q, precision, capi.KERNEL_TILE)
        assert rb == 0 and rt == 0
        np.testing.assert_array_equal(vb, vt)
        np.testing.assert_array_equal(cb, ct)
    m = 4000 if which == "fewest" else 1000
    rc, ov, oc, _, _ = O.ir_interpolate_at(t, Q[:m], nthreads=8)
    assert rc == 0
    assert np.max(np.abs(vb[:m] - ov)) <= 1e-12 * np.max(np.abs(ov))
    assert np.max(np.abs(cb[:m] - oc)) <= 1e-12 * np.max(np.abs(oc))


def test_binned_on_vertices_faces_and_edges():
    """Points on trellis vertices, cube edges and faces (weights exactly 0, so
    the slot-0 vertex is often dropped and the row table cannot serve the
    point) mixed with random points: binned equals the fused kernel bit for
    bit in FP32 and FP64, and FP64 matches the oracle."""
    g = S.grid(4, volume_fraction=1e-4)
    t = g.tables()
    dev = capi.Grid(t)
    verts = np.asarray(t["vertices"])
    cubes = np.asarray(t["cube_vertices"]).reshape(-1, 8)
    r = np.random.default_rng(31)
    pick = cubes[r.integers(0, len(cubes), 6000)]
    a, b = pick[:, 0], pick[:, r.integers(1, 8)]
    on_vertex = verts[a]
    on_segment = 0.5 * (verts[a] + verts[b])  # an edge, a face diagonal or the body diagonal
    B = np.asarray(t["outer_to_xyz"]).reshape(3, 3)
    special = np.linalg.solve(B, np.vstack([on_vertex, on_segment]).T).T
    Q = np.vstack([special, S.q_points(4, 50_000, stream=32)])
    Q = Q[r.permutation(len(Q))]
    say(f"binned on vertices/edges/faces: {len(special)} special of {len(Q)}")
    for precision in (capi.FP32, capi.FP64):
        rb, vb, cb = run_device(dev, Q, precision, capi.KERNEL_BINNED)
        rt, vt, ct = run_device(dev, Q, precision, capi.KERNEL_TILE)
        assert rb == 0 and rt == 0
        np.testing.assert_array_equal(vb, vt)
        np.testing.assert_array_equal(cb, ct)
    m = 15_000
    rc, ov, oc, _, _ = O.ir_interpolate_at(t, Q[:m], nthreads=8)
    assert rc == 0
    np.testing.assert_allclose(vb[:m], ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max())
    np.testing.assert_allclose(cb[:m], oc, rtol=1e-12, atol=1e-12)


def test_wide_rows_on_tetrahedron_vertices_and_faces():
    """A 24-atom payload with points on the vertices, edges and faces of the
    polyhedron nodes' tetrahedra mixed with random points: many have a zero
    slot-0 weight and no row-table block, so the binned pipeline sums them in
    its second launch (the wave per (point, mode) kernel over the listed
    points) while the rest go atom-major; binned equals the fused kernel bit
    for bit in FP32 and FP64, and FP64 matches the oracle."""
    grids, Qr, _ = S.batch_grids(n_grids=6, v_range=(800, 2000), n_q=20_000)
    sc = [int(sum(g.tables()["vectors_elements"])) for g in grids]
    t = grids[int(np.argmax(sc))].tables()
    dev = capi.Grid(t)
    verts = np.asarray(t["vertices"])
    tets = np.asarray(t["tet_vertices"]).reshape(-1, 4)
    r = np.random.default_rng(37)
    pick = tets[r.integers(0, len(tets), 3000)]
    i, j, k = pick[:, 0], pick[:, 1], pick[:, 2]
    pts = np.vstack([verts[i], 0.5 * (verts[i] + verts[j]), (verts[i] + verts[j] + verts[k]) / 3.0])
    B = np.asarray(t["outer_to_xyz"]).reshape(3, 3)
    special = np.linalg.solve(B, pts.T).T
    Q = np.vstack([special, Qr[:10_000]])
    Q = Q[r.permutation(len(Q))]
    say(f"wide rows on tetrahedra: {len(special)} special of {len(Q)}, {sc[int(np.argmax(sc))]} complex per mode")
    for precision in (capi.FP32, capi.FP64):
        rb, vb, cb = run_device(dev, Q, precision, capi.KERNEL_BINNED)
        rt, vt, ct = run_device(dev, Q, precision, capi.KERNEL_TILE)
        assert rb == 0 and rt == 0
        np.testing.assert_array_equal(vb, vt)
        np.testing.assert_array_equal(cb, ct)
    m = 1500
    rc, ov, oc, _, _ = O.ir_interpolate_at(t, Q[:m], nthreads=8)
    assert rc == 0
    assert np.max(np.abs(vb[:m] - ov)) <= 1e-12 * np.max(np.abs(ov))
    assert np.max(np.abs(cb[:m] - oc)) <= 1e-12 * np.max(np.abs(oc))
