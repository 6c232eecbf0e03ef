"""The learned tetrahedron proposal on the GPU, through the C ABI.

* the forward/backward kernels' loss and gradients against the CPU oracle
  (oracle/proposal.py, itself pinned by finite differences);
* the matrix-core inference against the oracle's bf16 emulation;
* training on pairs from the exact locator, then the binned pipeline with the
  proposal on and off: tetrahedra, indices, weights and outputs bit-identical,
  with the hit rate and the reduction of containment tests reported.
Needs an MI355X.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from oracle import oracle as O
from oracle import proposal as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def coarse():
    g = S.grid(1)
    t = g.tables()
    return t, capi.Grid(t)


def _random_batch(t, n, seed):
    r = np.random.default_rng(seed)
    off = np.asarray(t["poly_tet_offset"], np.int64)
    K = int(off[-1])
    x = r.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    labels = r.integers(0, K, size=n).astype(np.int32)
    return x, labels, off, K


def test_gradients_match_the_bf16_oracle(coarse):
    t, dev = coarse
    lib = capi.lib()
    x, labels, off, K = _random_batch(t, 300, 1)
    say(f"proposal gradients: K={K} tetrahedra, 300 samples")
    r = np.random.default_rng(2)
    params = (r.normal(size=P.n_params(K)) * 0.3).astype(np.float32)
    capi.check(lib.brille_proposal_set_params(dev.h, params.ctypes.data_as(C.POINTER(C.c_float)), params.size))
    loss = C.c_float()
    grads = np.zeros(params.size, np.float32)
    npar = C.c_uint32()
    capi.check(lib.brille_proposal_loss_grad(dev.h, x.ctypes.data_as(C.POINTER(C.c_float)),
                                             labels.ctypes.data_as(C.POINTER(C.c_int32)), len(labels),
                                             C.byref(loss), grads.ctypes.data_as(C.POINTER(C.c_float)), C.byref(npar)))
    assert npar.value == params.size
    # the training products run in bf16 on the matrix cores: against the
    # oracle's emulation of exactly those roundings the difference is FP32
    # summation order (and a ReLU mask flipped by it in rare entries)
    ref_loss, ref_g = P.loss_grad_bf16(params.astype(np.float64), x, labels, off, K)
    assert abs(loss.value - ref_loss) <= 1e-4 * max(1.0, abs(ref_loss))
    scale = np.abs(ref_g).max()
    np.testing.assert_allclose(grads, ref_g, rtol=2e-3, atol=2e-5 * scale)
    # and within bf16 resolution of the exact float64 gradient
    ex_loss, ex_g = P.loss_grad(params.astype(np.float64), x, labels, off, K)
    assert abs(loss.value - ex_loss) <= 2e-2 * max(1.0, abs(ex_loss))
    assert np.abs(grads - ex_g).max() <= 3e-2 * scale
    # parameters round-trip through the device
    back = np.zeros_like(params)
    capi.check(lib.brille_proposal_get_params(dev.h, back.ctypes.data_as(C.POINTER(C.c_float)), C.byref(npar)))
    np.testing.assert_array_equal(back, params)


def test_inference_matches_oracle(coarse):
    t, dev = coarse
    lib = capi.lib()
    x, labels, off, K = _random_batch(t, 5000, 3)
    say("proposal inference vs bf16 oracle")
    r = np.random.default_rng(4)
    params = (r.normal(size=P.n_params(K)) * 0.3).astype(np.float32)
    capi.check(lib.brille_proposal_set_params(dev.h, params.ctypes.data_as(C.POINTER(C.c_float)), params.size))
    out = np.zeros(len(labels), np.int32)
    capi.check(lib.brille_proposal_predict(dev.h, x.ctypes.data_as(C.POINTER(C.c_float)),
                                           labels.ctypes.data_as(C.POINTER(C.c_int32)), len(labels),
                                           out.ctypes.data_as(C.POINTER(C.c_int32))))
    t0, t1 = P.candidates(off, labels)
    assert np.all((out >= t0) & (out < t1))
    ref, gap = P.predict_bf16(params, x, t0, t1, K)
    differ = out != ref
    # only near ties may resolve differently (FP32 summation order)
    assert np.all(gap[differ] < 1e-3 * np.maximum(1.0, np.abs(gap[differ]))), (differ.sum(), gap[differ][:5])
    assert differ.mean() < 0.01


def _binned(dev, Q, use_proposal):
    lib = capi.lib()
    n = Q.shape[0]
    vs, cs = dev.out_shapes(n)
    vals = np.zeros(vs, np.float32)
    vecs = np.zeros(cs, np.complex64)
    bufs = dict(tet=np.zeros(n, np.uint32), n_iw=np.zeros(n, np.uint8), iw_index=np.zeros((n, 8), np.uint32),
                iw_weight=np.zeros((n, 8)), node=np.zeros(n, np.uint32), status=np.zeros(n, np.uint8))
    aux = capi.AuxOut(**{k: v.ctypes.data_as(C.c_void_p) for k, v in bufs.items()})
    o = capi.RunOpts(0, capi.FP32, int(use_proposal), capi.MEM_HOST, capi.KERNEL_BINNED, 1)
    capi.check(lib.brille_ir_interpolate_at(dev.h, Q.ctypes.data_as(C.c_void_p), n, C.byref(o),
                                            vals.ctypes.data_as(C.c_void_p), vecs.ctypes.data_as(C.c_void_p),
                                            C.byref(aux), None))
    return vals, vecs, bufs


def test_trained_proposal_keeps_the_reference_choice():
    g = S.grid(4, volume_fraction=1e-4)
    t = g.tables()
    dev = capi.Grid(t)
    lib = capi.lib()
    say(f"training the proposal: {t['tet_volume'].size} tetrahedra")
    opts = capi.TrainOpts(2_000_000, 7, 8, 8192, 2e-3)
    rep = capi.TrainReport()
    capi.check(lib.brille_proposal_train(dev.h, C.byref(opts), C.byref(rep)))
    say(f"trained on {rep.n_pairs} pairs in {rep.ms:.0f} ms: loss {rep.final_loss:.4f}, "
        f"held-out hit rate {rep.train_hit_rate:.3f}, K={rep.n_classes}")
    assert rep.n_pairs == 2_000_000 and rep.n_classes == t["tet_volume"].size
    assert np.isfinite(rep.final_loss) and rep.train_hit_rate > 0.5
    # random points plus points on and around tetrahedron faces and vertices
    Q = S.q_points(4, 1_000_000, stream=21)
    verts = np.asarray(t["vertices"])
    tv = np.asarray(t["tet_vertices"]).reshape(-1, 4)
    r = np.random.default_rng(5)
    pick = r.integers(0, len(tv), 20000)
    w = r.dirichlet(np.ones(3), size=len(pick))
    face = (w[:, :, None] * verts[tv[pick, :3]]).sum(axis=1)
    B = np.asarray(t["outer_to_xyz"]).reshape(3, 3)
    special = np.linalg.solve(B, np.vstack([face, verts[tv[pick[:2000], 0]]]).T).T
    Q = np.vstack([Q, special, np.nextafter(special, np.inf)])
    v0, c0, a0 = _binned(dev, Q, False)
    v1, c1, a1 = _binned(dev, Q, True)
    tried, hits = C.c_uint64(), C.c_uint64()
    capi.check(lib.brille_proposal_stats(dev.h, C.byref(tried), C.byref(hits)))
    for k in a0:
        np.testing.assert_array_equal(a1[k], a0[k], err_msg=k)
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(c1, c0)
    poly = a0["tet"] != 0xFFFFFFFF
    # the network runs for points in polyhedron nodes of at least 8 tetrahedra
    off = np.asarray(t["poly_tet_offset"], np.int64)
    pidx = np.asarray(t["node_index"], np.int64)[a0["node"][poly].astype(np.int64)]
    eligible = (off[pidx + 1] - off[pidx]) >= 8
    say(f"{poly.sum()} surface points, {eligible.sum()} in nodes of 8+ tetrahedra: {tried.value} proposals, "
        f"{hits.value} taken ({hits.value / max(tried.value, 1):.3f})")
    assert tried.value == eligible.sum()
    assert hits.value > 0.5 * tried.value
    # what the network saves: containment tests of the exact scan (position of the
    # reference's tetrahedron in its node) against one test per verified proposal
    tet = a0["tet"][poly].astype(np.int64)
    pos = (tet - off[pidx] + 1).astype(np.float64)
    rate = hits.value / tried.value
    with_net = np.where(eligible, rate * 1.0 + (1 - rate) * (1.0 + pos), pos)
    say(f"containment tests per surface point: exact scan {pos.mean():.2f}, with the proposal {with_net.mean():.2f}")
    # the same tetrahedra also against the oracle on a subsample
    m = 20000
    rc, ov, oc, oaux, _ = O.ir_interpolate_at(t, Q[:m], nthreads=8)
    assert rc == 0
    np.testing.assert_array_equal(a1["tet"][:m], oaux["tet"])


def _same_device_bytes(a, b, step=1 << 28):
    """Byte-wise equality of two device arrays, copied to the host in pieces."""
    lib = capi.lib()
    assert a.nbytes == b.nbytes
    ha = np.empty(step, np.uint8)
    hb = np.empty(step, np.uint8)
    for off in range(0, a.nbytes, step):
        k = min(step, a.nbytes - off)
        capi.check(lib.brille_memcpy_d2h(ha.ctypes.data_as(C.c_void_p), C.c_void_p(a.ptr.value + off), k))
        capi.check(lib.brille_memcpy_d2h(hb.ctypes.data_as(C.c_void_p), C.c_void_p(b.ptr.value + off), k))
        if not np.array_equal(ha[:k], hb[:k]):
            return False
    return True


def test_learned_proposal_over_several_passes():
    """Three passes of 2^24 points with the proposal on: each pass keeps its own
    list, hints and counter while the next is prepared; the outputs equal those
    of the binned pipeline without the proposal, byte for byte."""
    g = S.grid(4, volume_fraction=1e-4)
    t = g.tables()
    dev = capi.Grid(t)
    lib = capi.lib()
    opts = capi.TrainOpts(400_000, 9, 2, 8192, 2e-3)
    rep = capi.TrainReport()
    capi.check(lib.brille_proposal_train(dev.h, C.byref(opts), C.byref(rep)))
    n = (2 << 24) + 4099
    say(f"learned proposal over several passes: n={n}, held-out hit rate {rep.train_hit_rate:.3f}")
    Q = S.q_points(4, n, stream=23)
    dQ = capi.DeviceArray.from_host(Q)
    del Q
    vs, cs = dev.out_shapes(n)
    outs = []
    for use in (True, False):
        dv = capi.DeviceArray(vs, np.float32)
        dc = capi.DeviceArray(cs, np.complex64)
        o = capi.RunOpts(0, capi.FP32, int(use), capi.MEM_DEVICE, capi.KERNEL_BINNED, 1)
        capi.check(lib.brille_ir_interpolate_at(dev.h, dQ.ptr, n, C.byref(o), dv.ptr, dc.ptr, None, None))
        if use:
            tried, hits = C.c_uint64(), C.c_uint64()
            capi.check(lib.brille_proposal_stats(dev.h, C.byref(tried), C.byref(hits)))
            say(f"{tried.value} proposals, {hits.value} taken")
            assert tried.value > 0 and hits.value > 0.5 * tried.value
        outs.append((dv, dc))
    assert _same_device_bytes(outs[0][0], outs[1][0])
    assert _same_device_bytes(outs[0][1], outs[1][1])
