"""Configuration 2 at its stated size through one batch (needs an MI355X).

64 grids cycling cubic No. 221, tetragonal No. 139, orthorhombic No. 62,
monoclinic No. 14 and hexagonal No. 194, 1-24 atoms (3-72 modes), vertex
counts log-uniform in 2e3-2e5 (no payload cap: 65 GB of complex128
eigenvectors in all, generated grid by grid and filled keeping the FP32 rows
only), 1e7 Q split by Dirichlet(0.3) weights: 152 GB of FP32 results in one
autotuned, graph-replayed persistent launch.  The first (up to) 3000 points of
every grid are checked against the FP64 oracle: eigenvalues within 1e-5 of
the mode's value scale, eigenvectors within 1e-5 of the mode norm.
"""
import sys
import time

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CHECK = 3000


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_full_batch_matches_the_oracle():
    t0 = time.time()
    specs = S.batch_specs(payload_cap=None)
    Q, off = S.batch_queries()
    devs, want = [], []
    for spec in specs:
        i = spec["index"]
        g = S.batch_grid(spec)
        vals, vecs = S.batch_payload(g, spec)
        A = spec["atoms"]
        t = S.batch_layout_tables(g, spec)
        a, b = int(off[i]), int(off[i + 1])
        m = min(CHECK, b - a)
        rc, ov, oc, _, _ = O.ir_interpolate_at(t, Q[a:a + m], values=vals, vectors=vecs, nthreads=16)
        assert rc == 0
        dev = capi.Grid(t)
        dev.fill_ex(t, vals.ctypes.data, vecs.ctypes.data, capi.MEM_HOST, capi.FP64, capi.KEEP_FP32)
        devs.append(dev)
        want.append((ov, oc))
        say(f"grid {i}: sg {spec['sg']} A={A} V={vals.shape[0]} points {b - a} ({time.time() - t0:.0f} s)")
        del vals, vecs
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Q)
    vals, vecs = batch.outputs(off, capi.FP32)
    say(f"outputs: {sum(v.nbytes + c.nbytes for v, c in zip(vals, vecs)) / 1e9:.1f} GB on the device; autotuning")
    assert batch.run(dQ, off, vals, vecs, precision=capi.FP32, autotune=True) == 0
    say(f"schedule {batch.schedule()}")
    ms = batch.time(dQ, off, vals, vecs, precision=capi.FP32, iters=2)
    say(f"one graph replay: {ms:.1f} ms for {int(off[-1])} points")
    assert batch.run(dQ, off, vals, vecs, precision=capi.FP32) == 0
    worst_v = worst_c = 0.0
    for i, (ov, oc) in enumerate(want):
        m = ov.shape[0]
        if m == 0:
            continue
        gv = vals[i].rows_to_host(0, m).astype(np.float64)
        gc = vecs[i].rows_to_host(0, m).astype(np.complex128).reshape(oc.shape)
        scale = np.max(np.abs(ov), axis=0, keepdims=True)
        ev = np.max(np.abs(gv - ov) / scale)
        nrm = np.sqrt(np.sum(np.abs(oc.reshape(m, oc.shape[1], -1)) ** 2, axis=2))
        err = np.sqrt(np.sum(np.abs((gc - oc).reshape(m, oc.shape[1], -1)) ** 2, axis=2))
        ec = np.max(err / np.maximum(nrm, 1e-30))
        worst_v, worst_c = max(worst_v, ev), max(worst_c, ec)
        assert ev <= 1e-5 and ec <= 1e-5, (i, ev, ec)
    say(f"FP32 batch against the FP64 oracle: eigenvalues {worst_v:.2e}, eigenvectors {worst_c:.2e}")
