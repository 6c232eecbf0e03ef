"""brille_grid_sort on the GPU against the oracle's sort (needs an MI355X).

Payloads whose modes never cross, shuffled independently at every vertex
(synthetic.shuffled_modes), on the two small grids and on the benchmark's
121 200-vertex grid:
* keys and every key's permutation bit for bit against the oracle's
  DualInterpolator::sort restatement (src/interpolatordual.hpp:398-434);
* interpolation through the sorted table (TILE and BINNED, FP64) against the
  oracle's interpolation with the oracle's table: indices bit for bit,
  values and vectors within 1e-12;
* the sorted shuffled grid gives, point by point, exactly the unshuffled
  grid's results in the pivot vertex's mode order;
* BZTrellisQdc.fill(..., sort=True) runs the same sort.
"""
import sys
import time

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_table(t, nthreads=16):
    rc, keys, pairs, row, col = O.sorted_table(t, nthreads=nthreads)
    assert rc == 0
    V = np.uint64(np.asarray(t["vertices"]).shape[0])
    B = row.shape[1]
    fwd = {int(k): n for n, k in enumerate(pairs[:, 0].astype(np.uint64) * V + pairs[:, 1].astype(np.uint64))}
    rev = {int(k): n for n, k in enumerate(pairs[:, 1].astype(np.uint64) * V + pairs[:, 0].astype(np.uint64))}
    perms = np.zeros((len(keys), B), np.uint32)
    for n, k in enumerate(keys):
        k = int(k)
        perms[n] = row[fwd[k]] if k in fwd else col[rev[k]]
    return keys, perms


def device_table(dev):
    keys, index, perms = dev.permutations()
    assert (perms[0] == np.arange(perms.shape[1])).all()
    return keys, perms[index]


@pytest.fixture(scope="module", params=[0, 1])
def shuffled(request):
    cfg = request.param
    t = dict(S.grid(cfg).tables())
    values, vectors, order = S.shuffled_modes(t, seed=cfg)
    t["values"], t["vectors"] = values, vectors
    dev = capi.Grid(t)
    dev.sort()
    say(f"config {cfg}: sorted {len(dev.permutations()[0])} keys")
    return cfg, t, dev, order


def test_table_matches_the_oracle(shuffled):
    cfg, t, dev, order = shuffled
    ok, op = oracle_table(t)
    dk, dp = device_table(dev)
    np.testing.assert_array_equal(dk, ok)
    np.testing.assert_array_equal(dp, op)
    assert (dp != np.arange(dp.shape[1])).any(axis=1).mean() > 0.5  # the shuffles make most pairs non-trivial


@pytest.mark.parametrize("kernel", [capi.KERNEL_TILE, capi.KERNEL_BINNED])
def test_interpolation_through_the_sorted_table(shuffled, kernel):
    cfg, t, dev, order = shuffled
    keys, index, perms = dev.permutations()
    Q = S.q_points(cfg, 20_000, stream=61)
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=16, perm=(keys, index, perms))
    assert rc == 0
    aux = dict(tau=kernel != capi.KERNEL_BINNED)
    gv, gc, gaux = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=kernel, aux=aux)
    for k in gaux:
        np.testing.assert_array_equal(gaux[k], oaux[k], err_msg=k)
    assert np.max(np.abs(gv - ov)) <= 1e-12 * np.abs(ov).max()
    assert np.max(np.abs(gc - oc)) <= 1e-12 * np.abs(oc).max()


def test_sorting_restores_the_unshuffled_modes(shuffled):
    cfg, t, dev, order = shuffled
    rows = np.arange(order.shape[0])[:, None]
    inv = np.argsort(order, axis=1)
    t0 = dict(t)
    t0["values"], t0["vectors"] = t["values"][rows, inv], t["vectors"][rows, inv]
    ref = capi.Grid(t0)
    Q = S.q_points(cfg, 20_000, stream=62)
    for kernel in (capi.KERNEL_TILE, capi.KERNEL_BINNED):
        sv, sc, saux = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=kernel, aux=dict(tau=False))
        uv, uc = ref.ir_interpolate_at(Q, precision=capi.FP64, kernel=kernel)
        pivot = saux["iw_index"][:, 0]
        o = order[pivot]  # original mode of each output branch
        np.testing.assert_array_equal(sv, np.take_along_axis(uv, o.reshape(o.shape + (1,) * (uv.ndim - 2)), axis=1))
        np.testing.assert_array_equal(sc, np.take_along_axis(uc, o.reshape(o.shape + (1,) * (uc.ndim - 2)), axis=1))


def test_python_fill_with_sort(shuffled):
    cfg, t, dev, order = shuffled
    g = S.grid(cfg)
    g.fill(t["values"], np.asarray(t["values_elements"]), t["vectors"], np.asarray(t["vectors_elements"]), True)
    keys, index, perms = g.permutation_table()
    dk, dp = device_table(dev)
    np.testing.assert_array_equal(keys, dk)
    np.testing.assert_array_equal(perms[index], dp)


def test_sort_on_the_benchmark_grid():
    t = dict(S.grid(4).tables())
    values, vectors, order = S.shuffled_modes(t, seed=4)
    t["values"], t["vectors"] = values, vectors
    dev = capi.Grid(t)
    t0 = time.perf_counter()
    dev.sort()
    ms = (time.perf_counter() - t0) * 1e3
    dk, dp = device_table(dev)
    say(f"config 4 grid: {len(dk)} keys sorted in {ms:.1f} ms (first call, host-timed)")
    t1 = time.perf_counter()
    ok, op = oracle_table(t)
    say(f"  oracle: {(time.perf_counter() - t1):.1f} s on 16 threads")
    np.testing.assert_array_equal(dk, ok)
    np.testing.assert_array_equal(dp, op)
