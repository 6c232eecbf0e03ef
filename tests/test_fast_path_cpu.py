"""The fast fold-in and O(1) binning take exactly the exact path's decisions.

The kernels' own routines are compiled for the host (tests/harness/fold_check.hip)
and run side by side on the CPU: q, tau, Ridx, invRidx, status and the located node
must be bit-identical between the exact and the fast variant, on random Q and on
points placed on zone faces, edges, vertices, wedge planes and knot planes (and one
ulp either side of them).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from brille_amd import capi
from brille_amd import synthetic as S
from conftest import nacl_grid

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "harness", "fold_check.hip")
LIB = os.path.join(HERE, "harness", "libfold_check.so")


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(p) for p in [SRC] + [os.path.join(HERE, "..", "brille_amd", "csrc", f) for f in
                                                  ("exact_path.hpp", "fast_fold.hpp", "fast_tables.hpp",
                                                   "device_common.hpp")]):
        cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
               "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-o", LIB, SRC]
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert p.returncode == 0, p.stdout
    L = C.CDLL(LIB)
    L.fold_check.restype = C.c_int
    L.fold_check.argtypes = [C.POINTER(capi.GridDesc), C.c_void_p, C.c_uint64] + [C.c_void_p] * 7
    return L


def run(L, tables, Q):
    th = capi.TableHolder(tables)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    n = len(Q)
    q = np.zeros((2 * n, 3))
    tau = np.zeros((2 * n, 3), np.int32)
    r = np.zeros(2 * n, np.uint32)
    ir = np.zeros(2 * n, np.uint32)
    st = np.zeros(2 * n, np.uint8)
    nodes = np.zeros(2 * n, np.uint64)
    counts = np.zeros(2, np.uint64)
    L.fold_check(C.byref(th.desc), Q.ctypes.data, n, *[a.ctypes.data for a in (q, tau, r, ir, st, nodes, counts)])
    return q, tau, r, ir, st, nodes, counts


def special_points(g, cfg, rng):
    bz = g.BrillouinZone
    t = g.tables()
    B = t["outer_to_xyz"].reshape(3, 3)
    Binv = np.linalg.inv(B)
    ir = np.asarray(bz.ir_vertices)
    sp = [np.zeros((1, 3)), np.asarray(bz.vertices), np.asarray(bz.points), ir,
          (ir[:, None] + ir[None, :]).reshape(-1, 3) / 2,
          (np.asarray(bz.vertices)[:, None] + np.asarray(bz.vertices)[None, :]).reshape(-1, 3) / 2]
    inside = rng.dirichlet(np.ones(len(ir)), size=300) @ ir
    x = inside @ B.T
    for d in range(3):
        k = t["knots"][d]
        s = x.copy()
        s[:, d] = k[np.clip(np.searchsorted(k, s[:, d]), 0, len(k) - 1)]
        sp.append(s @ Binv.T)
    sp = np.vstack(sp)
    Rs = t["rotations"].reshape(-1, 3, 3)
    out = [sp]
    for j in range(len(Rs)):
        out.append(sp @ Rs[j] + rng.integers(-3, 4, size=(1, 3)))
    sp = np.vstack(out)
    return np.vstack([sp, np.nextafter(sp, np.inf), np.nextafter(sp, -np.inf)])


@pytest.mark.parametrize("cfg,vf", [(0, None), (1, None), (4, 1e-4)])
def test_fast_equals_exact(harness, cfg, vf):
    g = S.grid(cfg, volume_fraction=vf)
    t = g.tables()
    rng = np.random.default_rng(cfg)
    Q = np.vstack([S.q_points(cfg, 100000, stream=21), special_points(g, cfg, rng)])
    q, tau, r, ir, st, nodes, counts = run(harness, t, Q)
    assert counts[0] == 0, f"{counts[0]} of {len(Q)} points differ"
    assert np.array_equal(q[0::2], q[1::2])
    assert np.array_equal(r[0::2], r[1::2]) and np.array_equal(ir[0::2], ir[1::2])
    # the random points are decided almost entirely by the fast bounds
    assert counts[1] < 0.05 * len(Q)


def test_fast_equals_exact_nacl(harness, nacl):
    lat, bz, g, perm, q_nu, _ = nacl_grid(nacl)
    t = g.tables()
    rng = np.random.default_rng(3)
    Q = np.vstack([q_nu, rng.uniform(-2, 2, size=(50000, 3)), special_points(g, 0, rng)])
    q, tau, r, ir, st, nodes, counts = run(harness, t, Q)
    assert counts[0] == 0


def _near_boundaries(g, rng, scales):
    """Zone vertices, face points and wedge-plane points moved just outside and inside by each
    relative amount in `scales`, and their symmetry images shifted by lattice vectors."""
    bz = g.BrillouinZone
    t = g.tables()
    verts = np.asarray(bz.vertices)
    faces = np.asarray(bz.points)
    ir = np.asarray(bz.ir_vertices)
    mids = (ir[:, None] + ir[None, :]).reshape(-1, 3) / 2
    base = np.vstack([verts, faces, (verts[:, None] + verts[None, :]).reshape(-1, 3) / 2, mids])
    out = []
    for s in scales:
        out += [base * (1 + s), base * (1 - s)]
        d = rng.normal(size=base.shape)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        out.append(base + s * d)
    pts = np.vstack(out)
    Rs = t["rotations"].reshape(-1, 3, 3)
    imgs = [pts] + [pts @ Rs[j] + rng.integers(-2, 3, size=(1, 3)) for j in range(0, len(Rs), max(1, len(Rs) // 8))]
    return np.vstack(imgs)


@pytest.mark.parametrize("tol", [1e-10, 1e-8, 1e-6, 1e-4])
@pytest.mark.parametrize("cfg", [0, 1])
def test_fast_equals_exact_with_a_loose_zone_tolerance(harness, cfg, tol):
    """ApproxConfig's reciprocal-space tolerance widens the exact tests' zero band; the fast
    margins follow it, so points just outside faces and wedge planes are still decided alike."""
    g = S.grid(cfg)
    t = dict(g.tables())
    t["float_tolerance"] = tol
    rng = np.random.default_rng(int(-np.log10(tol)) + 10 * cfg)
    scales = np.concatenate([tol * np.array([0.3, 0.6, 0.9, 1.0, 1.1, 1.5, 2, 3, 10]), np.logspace(-12, -3, 10)])
    Q = np.vstack([S.q_points(cfg, 20000, stream=23), _near_boundaries(g, rng, scales)])
    q, tau, r, ir, st, nodes, counts = run(harness, t, Q)
    assert counts[0] == 0, f"{counts[0]} of {len(Q)} points differ at tolerance {tol}"
