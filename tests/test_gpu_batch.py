"""Batched persistent launch over many uneven grids (configuration 2 in small),
through the C ABI.  Needs an MI355X.

The graph-replayed batch must give what one call per grid gives; replays are
repeated with the program's own eager work and synchronisations between them
and every replay is checked; FP64 batch output matches the oracle.
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


@pytest.fixture(scope="module")
def small_batch():
    say("building 10 small grids of the batch configuration")
    grids, Q, offsets = S.batch_grids(n_grids=10, v_range=(500, 5000), n_q=200_000)
    tables = [g.tables() for g in grids]
    devs = [capi.Grid(t) for t in tables]
    return grids, tables, devs, Q, offsets


def per_grid(devs, dQ_host, offsets, precision, kernel=capi.KERNEL_TILE):
    out = []
    for g, dev in enumerate(devs):
        a, b = int(offsets[g]), int(offsets[g + 1])
        if b == a:
            out.append((None, None))
            continue
        v, c = dev.ir_interpolate_at(dQ_host[a:b], precision=precision, kernel=kernel)
        out.append((v, c))
    return out


@pytest.mark.parametrize("precision", [capi.FP32, capi.FP64])
def test_batch_equals_per_grid_calls(small_batch, precision):
    grids, tables, devs, Q, offsets = small_batch
    say(f"batch vs per-grid precision={precision}")
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Q)
    vals, vecs = batch.outputs(offsets, precision)
    assert batch.run(dQ, offsets, vals, vecs, precision=precision) == 0
    ref = per_grid(devs, Q, offsets, precision)
    for g in range(len(devs)):
        if ref[g][0] is None:
            continue
        np.testing.assert_array_equal(vals[g].to_host(), ref[g][0])
        np.testing.assert_array_equal(vecs[g].to_host(), ref[g][1])


def test_batch_graph_replays(small_batch):
    """Replays of the captured graph, with eager calls and synchronisations in
    between, each equal to the first result."""
    grids, tables, devs, Q, offsets = small_batch
    say("batch graph replays")
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Q)
    vals, vecs = batch.outputs(offsets, capi.FP32)
    assert batch.run(dQ, offsets, vals, vecs, precision=capi.FP32) == 0
    first = [(v.to_host(), c.to_host()) for v, c in zip(vals, vecs)]
    for rep in range(3):
        # eager work between replays: per-grid calls on the same handles
        per_grid(devs[:3], Q, offsets, capi.FP32)
        for v, c in zip(vals, vecs):  # clobber the outputs
            capi.check(capi.lib().brille_memcpy_h2d(v.ptr, np.full(v.shape, 7, v.dtype).ctypes.data_as(C.c_void_p), v.nbytes))
        capi.check(capi.lib().brille_device_sync())
        assert batch.run(dQ, offsets, vals, vecs, precision=capi.FP32) == 0
        for g, (v, c) in enumerate(zip(vals, vecs)):
            np.testing.assert_array_equal(v.to_host(), first[g][0], err_msg=f"replay {rep} grid {g}")
            np.testing.assert_array_equal(c.to_host(), first[g][1], err_msg=f"replay {rep} grid {g}")


def test_batch_autotune_and_sequential_baseline(small_batch):
    grids, tables, devs, Q, offsets = small_batch
    say("batch autotune")
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Q)
    vals, vecs = batch.outputs(offsets, capi.FP32)
    assert batch.run(dQ, offsets, vals, vecs, precision=capi.FP32, autotune=True) == 0
    sch = batch.schedule()
    assert sch["tile_points"] in (128, 256) and sch["blocks"] >= 1 and sch["tiles_per_grab"] >= 1
    tuned = [(v.to_host(), c.to_host()) for v, c in zip(vals, vecs)]
    # the unbatched baseline (one tile-kernel call per grid) through the same entry point
    v2, c2 = batch.outputs(offsets, capi.FP32)
    assert batch.run(dQ, offsets, v2, c2, precision=capi.FP32, kernel=capi.KERNEL_TILE) == 0
    for g in range(len(devs)):
        np.testing.assert_array_equal(v2[g].to_host(), tuned[g][0])
        np.testing.assert_array_equal(c2[g].to_host(), tuned[g][1])
    ms_batch = batch.time(dQ, offsets, vals, vecs, precision=capi.FP32)
    ms_seq = batch.time(dQ, offsets, v2, c2, precision=capi.FP32, kernel=capi.KERNEL_TILE)
    say(f"batch {ms_batch:.3f} ms, per-grid calls {ms_seq:.3f} ms, schedule {sch}")


def test_batch_with_binned_grids_equals_tile_calls(small_batch):
    grids, tables, devs, Q, offsets = small_batch
    say("batch with binned grids")
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Q)
    counts = np.diff(offsets)
    cut = int(np.median(counts[counts > 0]))
    batch.set_schedule(128, 4, 1, 1, binned_min=cut)  # the larger half of the grids binned
    assert batch.schedule()["binned_min"] == cut
    for prec in (capi.FP32, capi.FP64):
        vals, vecs = batch.outputs(offsets, prec)
        assert batch.run(dQ, offsets, vals, vecs, precision=prec) == 0
        v2, c2 = batch.outputs(offsets, prec)
        assert batch.run(dQ, offsets, v2, c2, precision=prec, kernel=capi.KERNEL_TILE) == 0
        for g in range(len(devs)):
            np.testing.assert_array_equal(vals[g].to_host(), v2[g].to_host(), err_msg=f"grid {g}")
            np.testing.assert_array_equal(vecs[g].to_host(), c2[g].to_host(), err_msg=f"grid {g}")
    with pytest.raises(capi.BrilleError):
        batch.set_schedule(100, 4, 1, 1)


def test_batch_fp64_matches_oracle(small_batch):
    grids, tables, devs, Q, offsets = small_batch
    say("batch fp64 vs oracle")
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Q)
    vals, vecs = batch.outputs(offsets, capi.FP64)
    assert batch.run(dQ, offsets, vals, vecs, precision=capi.FP64) == 0
    checked = 0
    for g in range(len(devs)):
        a, b = int(offsets[g]), int(offsets[g + 1])
        m = min(b - a, 3000)
        if m == 0:
            continue
        rc, ov, oc, _, _ = O.ir_interpolate_at(tables[g], Q[a:a + m], nthreads=8)
        assert rc == 0
        gv = vals[g].to_host()[:m]
        gc = vecs[g].to_host()[:m]
        np.testing.assert_allclose(gv, ov, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ov).max()))
        np.testing.assert_allclose(gc, oc, rtol=1e-12, atol=1e-12)
        checked += 1
    assert checked >= 5


def test_batch_reports_failing_grid(small_batch):
    grids, tables, devs, Q, offsets = small_batch
    say("batch error attribution")
    Qb = Q.copy()
    g = next(i for i in range(len(devs)) if offsets[i + 1] - offsets[i] > 10)
    Qb[int(offsets[g]) + 5] = [np.nan, 0.0, 0.0]
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Qb)
    vals, vecs = batch.outputs(offsets, capi.FP32)
    rc = batch.run(dQ, offsets, vals, vecs, precision=capi.FP32)
    assert rc != 0
    buf = C.create_string_buffer(512)
    capi.lib().brille_last_error(buf, 512)
    msg = buf.value.decode()
    assert msg.startswith(f"grid {g}:"), msg
    assert np.all(vals[g].to_host()[5] == 0)


def test_batch_follows_a_refilled_member(small_batch):
    """brille_grid_fill replaces a member's payload buffers: the next batch run describes that
    grid again (its fill generation changed) and gives what a per-grid call on the new payload
    gives.  A fill that fails its checks leaves the previous payload in place."""
    grids, tables, devs, Q, offsets = small_batch
    say("batch after a member is refilled")
    t0 = dict(tables[0])
    d0 = capi.Grid(t0)
    batch = capi.Batch([d0, devs[1]])
    off = np.array([0, 3000, 5000], np.uint64)
    Qs = Q[:5000]
    dQ = capi.DeviceArray.from_host(Qs)
    vals, vecs = batch.outputs(off, capi.FP32)
    assert batch.run(dQ, off, vals, vecs, precision=capi.FP32) == 0
    before = vals[0].to_host()
    # refill grid 0 with scaled values and conjugated vectors
    t1 = dict(t0)
    t1["values"] = 2.5 * np.asarray(t0["values"])
    t1["vectors"] = np.conj(np.asarray(t0["vectors"]))
    d0.fill(t1)
    assert batch.run(dQ, off, vals, vecs, precision=capi.FP32) == 0
    v, c = d0.ir_interpolate_at(Qs[:3000], precision=capi.FP32, kernel=capi.KERNEL_TILE)
    np.testing.assert_array_equal(vals[0].to_host(), v)
    np.testing.assert_array_equal(vecs[0].to_host(), c)
    assert not np.array_equal(before, v)
    # a fill that fails its checks changes nothing
    bad = dict(t1)
    bad["gamma_F0"] = np.asarray(t1["gamma_F0"]).copy()
    bad["gamma_F0"].flat[0] = 99
    with pytest.raises(capi.BrilleError, match="GammaTable"):
        d0.fill(bad)
    assert batch.run(dQ, off, vals, vecs, precision=capi.FP32) == 0
    np.testing.assert_array_equal(vals[0].to_host(), v)
    np.testing.assert_array_equal(vecs[0].to_host(), c)
