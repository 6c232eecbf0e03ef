"""Oracle parity on the benchmark's own grid (needs an MI355X).

Configuration 4 at its stated resolution (node volume 1e-5 of the irreducible
zone: 121 200 vertices, 90 381 tetrahedra), the grid the headline throughput is
measured on.  2e5 random Q uniform in [-4, 4]^3 plus the adversarial set (zone
faces, edges and vertices, wedge planes, knot planes, Gamma, their symmetry
images and one-ulp nudges) and points on tetrahedron faces and vertices:

* TILE in FP64: every aux output (q, tau, Ridx, invRidx, node, tetrahedron,
  weights, status) bit for bit, values and eigenvectors within 1e-12;
* TILE and BINNED in FP32: indices bit for bit, values within 1e-5 of the mode
  scale, eigenvectors within 1e-5 of the mode norm;
* BINNED with a trained learned proposal: the same indices and the same FP32
  bits as BINNED without it.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O
from test_gpu_parity import EXACT, adversarial_points

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def headline():
    g = S.grid(4)  # node volume 1e-5 of the irreducible zone, the benchmark's grid
    t = g.tables()
    say(f"config 4 grid: {t['vertices'].shape[0]} vertices, {t['tet_volume'].size} tetrahedra")
    Q = adversarial_points(g, 200_000, 4, stream=31)
    verts = np.asarray(t["vertices"])
    tv = np.asarray(t["tet_vertices"]).reshape(-1, 4)
    r = np.random.default_rng(32)
    pick = r.integers(0, len(tv), 20_000)
    w = r.dirichlet(np.ones(3), size=len(pick))
    face = (w[:, :, None] * verts[tv[pick, :3]]).sum(axis=1)
    B = np.asarray(t["outer_to_xyz"]).reshape(3, 3)
    special = np.linalg.solve(B, np.vstack([face, verts[tv[pick[:4000], 0]]]).T).T
    Q = np.vstack([Q, special, np.nextafter(special, -np.inf)])
    say(f"{len(Q)} points; oracle ...")
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q, nthreads=8)
    assert rc == 0
    return t, capi.Grid(t), Q, ov, oc, oaux


def _fp32_close(gv, gc, ov, oc):
    scale = np.max(np.abs(ov), axis=0, keepdims=True)
    ev = np.max(np.abs(gv - ov) / scale)
    nrm = np.sqrt(np.sum(np.abs(oc.reshape(oc.shape[0], oc.shape[1], -1)) ** 2, axis=2))
    err = np.sqrt(np.sum(np.abs((gc - oc).reshape(oc.shape[0], oc.shape[1], -1)) ** 2, axis=2))
    ec = np.max(err / np.maximum(nrm, 1e-30))
    return ev, ec


def test_fp64_tile_is_bit_exact(headline):
    t, dev, Q, ov, oc, oaux = headline
    say("config 4 grid: TILE FP64")
    gv, gc, gaux = dev.ir_interpolate_at(Q, precision=capi.FP64, kernel=capi.KERNEL_TILE, aux=True)
    for k in EXACT:
        np.testing.assert_array_equal(gaux[k], oaux[k], err_msg=k)
    assert np.max(np.abs(gv - ov)) <= 1e-12 * np.abs(ov).max()
    assert np.max(np.abs(gc - oc)) <= 1e-12 * np.abs(oc).max()


@pytest.mark.parametrize("kernel", [capi.KERNEL_TILE, capi.KERNEL_BINNED])
def test_fp32_indices_exact_values_within_1e5(headline, kernel):
    t, dev, Q, ov, oc, oaux = headline
    say(f"config 4 grid: FP32 kernel {kernel}")
    aux = dict(tau=kernel != capi.KERNEL_BINNED)
    gv, gc, gaux = dev.ir_interpolate_at(Q, precision=capi.FP32, kernel=kernel, aux=aux)
    for k in gaux:
        np.testing.assert_array_equal(gaux[k], oaux[k], err_msg=k)
    ev, ec = _fp32_close(gv, gc, ov, oc)
    say(f"  eigenvalue error {ev:.2e} of the mode scale, eigenvector error {ec:.2e} of the mode norm")
    assert ev <= 1e-5 and ec <= 1e-5


def test_learned_proposal_keeps_the_oracle_indices(headline):
    t, dev, Q, ov, oc, oaux = headline
    lib = capi.lib()
    opts = capi.TrainOpts(4_000_000, 11, 2, 8192, 2e-3)
    rep = capi.TrainReport()
    capi.check(lib.brille_proposal_train(dev.h, C.byref(opts), C.byref(rep)))
    say(f"config 4 grid: proposal trained on {rep.n_pairs} pairs, held-out hit rate {rep.train_hit_rate:.3f}")
    aux = dict(tau=False)
    lv, lc, laux = dev.ir_interpolate_at(Q, precision=capi.FP32, kernel=capi.KERNEL_BINNED, aux=aux,
                                         use_proposal=True)
    tried, hits = C.c_uint64(), C.c_uint64()
    capi.check(lib.brille_proposal_stats(dev.h, C.byref(tried), C.byref(hits)))
    say(f"  {tried.value} proposals, {hits.value} verified (points on tetrahedron faces and vertices "
        f"are never taken from a proposal: the exact scan decides them)")
    assert tried.value > 1000 and hits.value > 1000
    for k in laux:
        np.testing.assert_array_equal(laux[k], oaux[k], err_msg=k)
    bv, bc = dev.ir_interpolate_at(Q, precision=capi.FP32, kernel=capi.KERNEL_BINNED)
    np.testing.assert_array_equal(lv, bv)
    np.testing.assert_array_equal(lc, bc)
    # on uniform Q most proposals are taken
    dev.ir_interpolate_at(S.q_points(4, 1_000_000, stream=33), precision=capi.FP32, kernel=capi.KERNEL_BINNED,
                          use_proposal=True)
    capi.check(lib.brille_proposal_stats(dev.h, C.byref(tried), C.byref(hits)))
    say(f"  uniform Q: {tried.value} proposals, {hits.value} verified")
    assert hits.value > 0.5 * tried.value
