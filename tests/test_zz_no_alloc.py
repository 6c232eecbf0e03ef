"""brille_ir_interpolate_at_noalloc against brille_ir_interpolate_at (needs an MI355X).

The allocation-free entry point runs the same kernels on a file-scope device
array instead of buffers the library allocates: its results must equal the
existing entry point's bit for bit (the pipelines differ in pass size and
stream use only, and every point is computed independently), on two calls in
a row with the second smaller, over several passes, with aux outputs, with
the learned proposal and on failing points.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _entry():
    return capi.lib().brille_ir_interpolate_at_noalloc


AUX = dict(ir_q=((3,), np.float64), tau=((3,), np.int32), rot=((), np.uint32), invrot=((), np.uint32),
           node=((), np.uint32), tet=((), np.uint32), n_iw=((), np.uint8), iw_index=((8,), np.uint32),
           iw_weight=((8,), np.float64), status=((), np.uint8))


def _call(dev, fn, dQ, n, precision, kernel, aux_names=(), use_proposal=False, check=True):
    vs, cs = dev.out_shapes(n)
    dv = capi.DeviceArray(vs, np.float32 if precision == capi.FP32 else np.float64)
    dc = capi.DeviceArray(cs, np.complex64 if precision == capi.FP32 else np.complex128)
    bufs = {k: capi.DeviceArray((n,) + AUX[k][0], AUX[k][1]) for k in aux_names}
    a = capi.AuxOut(**{k: v.ptr for k, v in bufs.items()}) if bufs else None
    o = capi.RunOpts(0, precision, int(use_proposal), capi.MEM_DEVICE, kernel, int(check))
    rc = fn(dev.h, dQ.ptr, n, C.byref(o), dv.ptr, dc.ptr, C.byref(a) if a is not None else None, None)
    capi.check(capi.lib().brille_device_sync())
    return rc, dv.to_host(), dc.to_host(), {k: v.to_host() for k, v in bufs.items()}


@pytest.fixture(scope="module")
def grid4():
    t = S.grid(4, volume_fraction=1e-4).tables()
    return t, capi.Grid(t)


@pytest.mark.parametrize("precision", [capi.FP32, capi.FP64])
@pytest.mark.parametrize("kernel", [capi.KERNEL_BINNED, capi.KERNEL_TILE, capi.KERNEL_AUTO])
def test_noalloc_equals_the_allocating_entry_point(grid4, precision, kernel):
    t, dev = grid4
    noalloc = _entry()
    regular = capi.lib().brille_ir_interpolate_at
    for n in (300_000, 70_001):  # two calls in a row, the second smaller
        say(f"no-alloc vs regular: precision={precision} kernel={kernel} n={n}")
        Q = S.q_points(4, n, stream=90 + n % 7)
        dQ = capi.DeviceArray.from_host(Q)
        aux = ("rot", "node", "tet", "iw_index", "iw_weight", "status")
        if kernel == capi.KERNEL_TILE:
            aux = aux + ("tau",)
        rc0, v0, c0, a0 = _call(dev, regular, dQ, n, precision, kernel, aux)
        rc1, v1, c1, a1 = _call(dev, noalloc, dQ, n, precision, kernel, aux)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
        for k in a0:
            assert np.array_equal(a0[k], a1[k]), k


def test_noalloc_over_several_passes(grid4):
    """1.2e6 points: three passes of 2^19 in the scratch array against the regular pipeline's one."""
    t, dev = grid4
    n = 1_200_000
    say(f"no-alloc over several passes: n={n}")
    Q = S.q_points(4, n, stream=95)
    dQ = capi.DeviceArray.from_host(Q)
    rc0, v0, c0, _ = _call(dev, capi.lib().brille_ir_interpolate_at, dQ, n, capi.FP32, capi.KERNEL_BINNED)
    rc1, v1, c1, _ = _call(dev, _entry(), dQ, n, capi.FP32, capi.KERNEL_BINNED)
    assert rc0 == 0 and rc1 == 0
    assert np.array_equal(v0, v1) and np.array_equal(c0, c1)


def test_noalloc_with_the_learned_proposal(grid4):
    t, dev = grid4
    lib = capi.lib()
    opts = capi.TrainOpts(300_000, 3, 1, 8192, 2e-3)
    rep = capi.TrainReport()
    capi.check(lib.brille_proposal_train(dev.h, C.byref(opts), C.byref(rep)))
    n = 600_000
    say(f"no-alloc with the proposal: n={n}")
    Q = S.q_points(4, n, stream=96)
    dQ = capi.DeviceArray.from_host(Q)
    rc0, v0, c0, _ = _call(dev, lib.brille_ir_interpolate_at, dQ, n, capi.FP32, capi.KERNEL_BINNED, use_proposal=True)
    tried, hits = C.c_uint64(), C.c_uint64()
    capi.check(lib.brille_proposal_stats(dev.h, C.byref(tried), C.byref(hits)))
    rc1, v1, c1, _ = _call(dev, _entry(), dQ, n, capi.FP32, capi.KERNEL_BINNED, use_proposal=True)
    tried1, hits1 = C.c_uint64(), C.c_uint64()
    capi.check(lib.brille_proposal_stats(dev.h, C.byref(tried1), C.byref(hits1)))
    assert rc0 == 0 and rc1 == 0
    assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    assert tried1.value == tried.value > 0 and hits1.value == hits.value


def test_noalloc_errors_and_arguments(grid4):
    t, dev = grid4
    fn = _entry()
    Q = S.q_points(4, 50_000, stream=97)
    Q[4321] = np.nan
    dQ = capi.DeviceArray.from_host(Q)
    for kernel in (capi.KERNEL_BINNED, capi.KERNEL_TILE):
        rc, *_ = _call(dev, fn, dQ, len(Q), capi.FP32, kernel)
        assert capi.STATUS_NAMES[rc] == "OUTSIDE_BZ"
        buf = C.create_string_buffer(256)
        capi.lib().brille_last_error(buf, 256)
        assert b"nan" in buf.value
        rc, v, c, a = _call(dev, fn, dQ, len(Q), capi.FP32, kernel, ("status",), check=False)
        assert rc == 0 and a["status"][4321] == 1 and (a["status"] == 0).sum() == len(Q) - 1
    # host memory is refused (it would need staging buffers)
    Qh = np.zeros((10, 3))
    o = capi.RunOpts(0, capi.FP32, 0, capi.MEM_HOST, capi.KERNEL_AUTO, 1)
    out = np.zeros((10, 6, 1), np.float32), np.zeros((10, 6, 6), np.complex64)
    rc = fn(dev.h, Qh.ctypes.data, 10, C.byref(o), out[0].ctypes.data, out[1].ctypes.data, None, None)
    assert capi.STATUS_NAMES[rc] == "INVALID_ARGUMENT"
