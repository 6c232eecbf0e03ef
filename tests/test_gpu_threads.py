"""Several host threads on one grid handle (needs an MI355X).

The reference's ir_interpolate_at is const and re-entrant
(src/bz_trellis.hpp:153-155); the Python binding here releases the GIL around
the device work, so threads really do overlap.  Every call holds the handle's
lock for its duration (error record, staging, work space, side stream and the
lazy device create and fill are per handle), and each thread must get exactly
what the same call gives when made alone.
"""
import sys
import threading

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run_threads(fns):
    out = [None] * len(fns)
    errs = []

    def work(i, f):
        try:
            out[i] = f()
        except Exception as e:  # noqa: BLE001 - reported below
            errs.append(e)

    ts = [threading.Thread(target=work, args=(i, f)) for i, f in enumerate(fns)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not errs, errs
    return out


def test_threads_on_one_pybind_grid_get_the_serial_results():
    say("threads on one BZTrellisQdc (lazy device create + fill in the race)")
    g = S.grid(1)  # not yet on the device: the first calls race to create and fill it
    cases = [(S.q_points(1, n, stream=60 + k), prec, kern)
             for k, (n, prec, kern) in enumerate([(300_000, "fp32", "binned"), (70_000, "fp64", "tile"),
                                                  (300_000, "fp64", "binned"), (5_000, "fp32", "plain"),
                                                  (200_000, "fp32", "auto"), (40_000, "fp64", "auto")])]
    fns = [lambda Q=Q, p=p, k=k: g.ir_interpolate_at(Q, precision=p, kernel=k) for Q, p, k in cases] * 2
    got = _run_threads(fns)
    for i, (Q, p, k) in enumerate(cases * 2):
        v, c = g.ir_interpolate_at(Q, precision=p, kernel=k)
        np.testing.assert_array_equal(got[i][0], v, err_msg=f"call {i}")
        np.testing.assert_array_equal(got[i][1], c, err_msg=f"call {i}")


def test_threads_on_one_c_abi_handle_with_errors_and_proposals():
    """ctypes releases the GIL during foreign calls: threads share one brille_grid, some of them
    with failing points (each must get its own error), one training the proposal meanwhile."""
    say("threads on one C-ABI handle")
    import ctypes as C
    t = S.grid(4, volume_fraction=1e-4).tables()
    dev = capi.Grid(t)
    good = [S.q_points(4, n, stream=70 + k) for k, n in enumerate([400_000, 50_000, 400_000])]
    bad = S.q_points(4, 20_000, stream=80)
    bad[1234] = np.nan

    def call_good(Q, prec, kern):
        return dev.ir_interpolate_at(Q, precision=prec, kernel=kern)

    def call_bad():
        try:
            dev.ir_interpolate_at(bad, precision=capi.FP32)
        except capi.BrilleError as e:
            return e.status, str(e)
        return None

    def train():
        opts = capi.TrainOpts(300_000, 5, 1, 8192, 2e-3)
        rep = capi.TrainReport()
        capi.check(capi.lib().brille_proposal_train(dev.h, C.byref(opts), C.byref(rep)))
        return rep.n_pairs

    fns = [lambda: call_good(good[0], capi.FP32, capi.KERNEL_BINNED), call_bad,
           lambda: call_good(good[1], capi.FP64, capi.KERNEL_TILE), train,
           lambda: call_good(good[2], capi.FP32, capi.KERNEL_AUTO), call_bad]
    got = _run_threads(fns)
    for i, (Q, prec, kern) in zip((0, 2, 4), [(good[0], capi.FP32, capi.KERNEL_BINNED),
                                             (good[1], capi.FP64, capi.KERNEL_TILE),
                                             (good[2], capi.FP32, capi.KERNEL_AUTO)]):
        v, c = dev.ir_interpolate_at(Q, precision=prec, kernel=kern)
        np.testing.assert_array_equal(got[i][0], v)
        np.testing.assert_array_equal(got[i][1], c)
    for i in (1, 5):
        assert got[i][0] == "OUTSIDE_BZ" and "nan" in got[i][1]
    assert got[3] == 300_000
