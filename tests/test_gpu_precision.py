"""FP32/complex64 sums against the FP64 oracle, per configuration (the
precision contract: indices bit-exact, eigenvalues and eigenvectors within
1e-5 of the mode scale), including configuration 3 (monoclinic P2_1/c,
24 atoms, 72 modes; general payload layout) at a reduced vertex count.
Needs an MI355X.
"""
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


def rel_errors(v, c, ov, oc):
    """Largest eigenvalue error over the mode's value scale and eigenvector
    error over the mode's vector norm."""
    vs = np.max(np.abs(ov), axis=0, keepdims=True)
    ev = float(np.max(np.abs(v - ov) / np.maximum(vs, 1e-300)))
    on = np.sqrt(np.sum(np.abs(oc.reshape(oc.shape[0], oc.shape[1], -1)) ** 2, axis=2, keepdims=True))
    d = np.abs(c - oc).reshape(oc.shape[0], oc.shape[1], -1)
    ec = float(np.max(d / np.maximum(on, 1e-300))) if oc.size else 0.0
    return ev, ec


CASES = [(0, None, 200_000), (1, None, 200_000), (3, 2e-3, 60_000), (4, 1e-4, 200_000)]


@pytest.mark.parametrize("cfg,vf,n", CASES)
@pytest.mark.parametrize("kernel", [capi.KERNEL_TILE, capi.KERNEL_BINNED])
def test_fp32_error_against_fp64_oracle(cfg, vf, n, kernel):
    say(f"fp32 error cfg{cfg} vf={vf} n={n} kernel={kernel}")
    g = S.grid(cfg, volume_fraction=vf)
    t = g.tables()
    dev = capi.Grid(t)
    Q = S.q_points(cfg, n, stream=31)
    out = dev.ir_interpolate_at(Q, precision=capi.FP32, kernel=kernel, aux=(kernel == capi.KERNEL_TILE))
    v32, c32 = out[0], out[1]
    a32 = out[2] if len(out) > 2 else None
    m = min(n, 5000 if cfg == 3 else 40_000)
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q[:m], nthreads=8)
    assert rc == 0
    if kernel == capi.KERNEL_TILE:
        for k in ("tau", "rot", "invrot", "node", "tet", "iw_index"):
            np.testing.assert_array_equal(a32[k][:m], oaux[k], err_msg=k)
    ev, ec = rel_errors(v32[:m].astype(np.float64), c32[:m].astype(np.complex128), ov, oc)
    say(f"cfg{cfg}: FP32 eigenvalue error {ev:.2e} of the mode scale, eigenvector error {ec:.2e} of the mode norm")
    assert ev <= 1e-5 and ec <= 1e-5


def test_config3_fp64_parity():
    """The general layout (72 modes, 24 atoms, Gamma rotation) in FP64 against the oracle."""
    say("cfg3 fp64 parity")
    g = S.grid(3, volume_fraction=2e-3)
    t = g.tables()
    dev = capi.Grid(t)
    Q = S.q_points(3, 5000, stream=32)
    v, c, a = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=capi.KERNEL_TUNED, aux=True)
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    assert rc == 0
    for k in ("ir_q", "tau", "rot", "invrot", "node", "tet", "n_iw", "iw_index", "iw_weight"):
        np.testing.assert_array_equal(a[k], oaux[k], err_msg=k)
    np.testing.assert_allclose(v, ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max())
    np.testing.assert_allclose(c, oc, rtol=1e-12, atol=1e-12)
