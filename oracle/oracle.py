"""ctypes front end of oracle/oracle.c -- TEST INFRASTRUCTURE ONLY.

The oracle consumes the same frozen tables (brille_grid_desc) the GPU path
consumes, so its outputs are directly comparable with the kernels'.
"""
import ctypes as C
import os

import numpy as np

from brille_amd.capi import AuxOut, CostInfo, DataHolder, GridDesc, DataDesc, GammaDesc, PermDesc, TableHolder, cost_info

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None
_dp = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from .build import build
            build()
        L = C.CDLL(LIB)
        L.brille_oracle_ir_interpolate_at.restype = C.c_int
        L.brille_oracle_ir_interpolate_at.argtypes = [
            C.POINTER(GridDesc), C.POINTER(DataDesc), C.POINTER(DataDesc), C.POINTER(GammaDesc), C.POINTER(PermDesc),
            _dp, C.c_uint64, C.c_int, C.c_int, _dp, _dp, C.POINTER(AuxOut), C.POINTER(C.c_uint64)]
        L.brille_oracle_ir_moveinto.restype = C.c_int
        L.brille_oracle_ir_moveinto.argtypes = [C.POINTER(GridDesc), _dp, C.c_uint64, _dp, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p]
        L.brille_oracle_moveinto.restype = C.c_int
        L.brille_oracle_moveinto.argtypes = [C.POINTER(GridDesc), _dp, C.c_uint64, _dp, C.c_void_p, C.c_void_p]
        L.brille_oracle_locate.restype = C.c_int
        L.brille_oracle_locate.argtypes = [C.POINTER(GridDesc), _dp, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, _dp]
        L.brille_oracle_orient3d_lattice.restype = C.c_double
        L.brille_oracle_orient3d_lattice.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_double]
        L.brille_oracle_orient3d_xyz.restype = C.c_double
        L.brille_oracle_orient3d_xyz.argtypes = [_dp, _dp, _dp, _dp]
        L.brille_oracle_approx.restype = C.c_int
        L.brille_oracle_approx.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.brille_oracle_find_bin.restype = C.c_uint32
        L.brille_oracle_find_bin.argtypes = [_dp, C.c_uint32, C.c_double]
        L.brille_oracle_stage_seconds.restype = None
        L.brille_oracle_stage_seconds.argtypes = [_dp]
        L.brille_oracle_lap.restype = C.c_int
        L.brille_oracle_lap.argtypes = [C.c_uint32, _dp, C.c_void_p, C.c_void_p]
        L.brille_oracle_sort_pairs.restype = C.c_int
        L.brille_oracle_sort_pairs.argtypes = [C.c_uint64, C.c_void_p, C.POINTER(DataDesc), C.POINTER(CostInfo),
                                               C.POINTER(DataDesc), C.POINTER(CostInfo), C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_int]
        L.brille_oracle_antiphase.restype = None
        L.brille_oracle_antiphase.argtypes = [C.c_uint32, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ir_interpolate_at(tables, Q, no_move=False, nthreads=1, values=None, vectors=None, perm=None, first_zone=False):
    """Returns (rc, values, vectors, aux, first_bad) computed on the CPU.
    first_zone: BrillouinZoneTrellis3::interpolate_at instead (moveinto only, no rotation)."""
    th = TableHolder(tables)
    dh = DataHolder(tables, values, vectors, perm)
    Q = _d(Q).reshape(-1, 3)
    n = Q.shape[0]
    vals = np.zeros((n,) + dh.vals.shape[1:], np.float64)
    vecs = np.zeros((n,) + dh.vecs.shape[1:], np.complex128)
    bufs = dict(ir_q=np.zeros((n, 3)), tau=np.zeros((n, 3), np.int32), rot=np.zeros(n, np.uint32),
                invrot=np.zeros(n, np.uint32), node=np.full(n, 0xFFFFFFFF, np.uint32),
                tet=np.full(n, 0xFFFFFFFF, np.uint32), n_iw=np.zeros(n, np.uint8),
                iw_index=np.zeros((n, 8), np.uint32), iw_weight=np.zeros((n, 8)), status=np.zeros(n, np.uint8))
    aux = AuxOut(**{k: v.ctypes.data_as(C.c_void_p) for k, v in bufs.items()})
    bad = C.c_uint64(0)
    rc = lib().brille_oracle_ir_interpolate_at(
        C.byref(th.desc), C.byref(dh.vdesc), C.byref(dh.cdesc), C.byref(dh.gdesc) if dh.gdesc is not None else None,
        C.byref(dh.pdesc) if dh.pdesc is not None else None, Q.ctypes.data_as(_dp), n,
        1 if no_move else (2 if first_zone else 0), int(nthreads),
        vals.ctypes.data_as(_dp), vecs.ctypes.data_as(_dp), C.byref(aux), C.byref(bad))
    return rc, vals, vecs, bufs, bad.value


def stage_seconds():
    """Wall seconds of the last ir_interpolate_at's stages on this thread:
    (a) fold-in, (b)+(c1) location and sums, (c2) rotation."""
    out = np.zeros(3)
    lib().brille_oracle_stage_seconds(out.ctypes.data_as(_dp))
    return out


def ir_moveinto(tables, Q):
    th = TableHolder(tables)
    Q = _d(Q).reshape(-1, 3)
    n = Q.shape[0]
    q = np.zeros((n, 3))
    tau = np.zeros((n, 3), np.int32)
    r = np.zeros(n, np.uint32)
    ir = np.zeros(n, np.uint32)
    st = np.zeros(n, np.uint8)
    rc = lib().brille_oracle_ir_moveinto(C.byref(th.desc), Q.ctypes.data_as(_dp), n, q.ctypes.data_as(_dp),
                                         tau.ctypes.data_as(C.c_void_p), r.ctypes.data_as(C.c_void_p),
                                         ir.ctypes.data_as(C.c_void_p), st.ctypes.data_as(C.c_void_p))
    return rc, q, tau, r, ir, st


def moveinto(tables, Q):
    th = TableHolder(tables)
    Q = _d(Q).reshape(-1, 3)
    n = Q.shape[0]
    q = np.zeros((n, 3))
    tau = np.zeros((n, 3), np.int32)
    st = np.zeros(n, np.uint8)
    rc = lib().brille_oracle_moveinto(C.byref(th.desc), Q.ctypes.data_as(_dp), n, q.ctypes.data_as(_dp),
                                      tau.ctypes.data_as(C.c_void_p), st.ctypes.data_as(C.c_void_p))
    return rc, q, tau, st


def get_xyz(tables, q):
    """Cartesian coordinates of rlu rows q as the reference forms them (LVec::get_xyz,
    src/array_lvec_methods.tpp:40-51, via utils::multiply_matrix_vector,
    src/utilities.tpp:44-55): x_i = (B_i0 q_0 + B_i1 q_1) + B_i2 q_2, each product
    and sum rounded on its own (numpy ufuncs; a matmul may sum in another order)."""
    B = np.asarray(tables["outer_to_xyz"], dtype=np.float64).reshape(3, 3)
    q = _d(q).reshape(-1, 3)
    x = np.empty_like(q)
    for i in range(3):
        x[:, i] = (B[i, 0] * q[:, 0] + B[i, 1] * q[:, 1]) + B[i, 2] * q[:, 2]
    return x


def locate(tables, xyz):
    th = TableHolder(tables)
    x = _d(xyz).reshape(-1, 3)
    n = x.shape[0]
    node = np.zeros(n, np.uint32)
    tet = np.zeros(n, np.uint32)
    niw = np.zeros(n, np.uint8)
    idx = np.zeros((n, 8), np.uint32)
    w = np.zeros((n, 8))
    rc = lib().brille_oracle_locate(C.byref(th.desc), x.ctypes.data_as(_dp), n, node.ctypes.data_as(C.c_void_p),
                                    tet.ctypes.data_as(C.c_void_p), niw.ctypes.data_as(C.c_void_p),
                                    idx.ctypes.data_as(C.c_void_p), w.ctypes.data_as(_dp))
    return rc, node, tet, niw, idx, w


def orient3d_lattice(a, b, c, d, Grecip, Greal, scale):
    args = [_d(x) for x in (a, b, c, d, Grecip, Greal)]
    return lib().brille_oracle_orient3d_lattice(*[x.ctypes.data_as(_dp) for x in args], float(scale))


def orient3d_xyz(a, b, c, d):
    args = [_d(x) for x in (a, b, c, d)]
    return lib().brille_oracle_orient3d_xyz(*[x.ctypes.data_as(_dp) for x in args])


def approx(a, b, Ttol=0.0, tol=1):
    return bool(lib().brille_oracle_approx(float(a), float(b), float(Ttol), int(tol)))


def find_bin(edges, x):
    e = _d(edges)
    return int(lib().brille_oracle_find_bin(e.ctypes.data_as(_dp), e.shape[0], float(x)))


def antiphase(a, b):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    out = np.zeros(2)
    lib().brille_oracle_antiphase(a.shape[0], a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return complex(out[0], out[1])


def lap(cost):
    """Exact linear assignment of a square cost matrix: (row, col) with
    row[i] the column of row i and col[j] the row of column j."""
    c = _d(cost)
    n = c.shape[0]
    row = np.zeros(n, np.int32)
    col = np.zeros(n, np.int32)
    rc = lib().brille_oracle_lap(n, c.ctypes.data_as(_dp), row.ctypes.data, col.ctypes.data)
    assert rc == 0
    return row, col


def sort_keys(tables):
    """The permutation-table keys of a trellis, i*V + j for every ordered pair
    of distinct vertices that share a cube node or a tetrahedron
    (PolyTrellis::collect_keys, src/trellis_poly_parallel.tpp:487-524;
    permutation_table_keys_from_indicies, src/permutation_table.hpp:284-292),
    sorted."""
    V = np.uint64(np.asarray(tables["vertices"]).shape[0])
    units = []
    cv = np.asarray(tables["cube_vertices"], dtype=np.uint64).reshape(-1, 8)
    if cv.size:
        units.append(cv)
    tv = np.asarray(tables["tet_vertices"], dtype=np.uint64).reshape(-1, 4)
    if tv.size:
        units.append(tv)
    keys = []
    for u in units:
        k = u.shape[1]
        for a in range(k):
            for b in range(k):
                if a == b:
                    continue
                i, j = u[:, a], u[:, b]
                sel = i != j
                keys.append(i[sel] * V + j[sel])
    if not keys:
        return np.zeros(0, np.uint64)
    return np.unique(np.concatenate(keys))


def sort_pairs(tables, pairs, values=None, vectors=None, values_cost=None, vectors_cost=None, nthreads=8,
               want_cost=False):
    """DualInterpolator cost_matrix + LAPJV for the (i, j) rows of `pairs`:
    (rc, row, col[, cost]) with row[k] the assignment stored for key (i, j) and
    col[k] the one for (j, i)."""
    dh = DataHolder(tables, values, vectors)
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    n = pairs.shape[0]
    B = dh.vdesc.branches
    row = np.zeros((n, B), np.int32)
    col = np.zeros((n, B), np.int32)
    cost = np.zeros((n, B, B)) if want_cost else None
    vc = values_cost if values_cost is not None else cost_info()
    cc = vectors_cost if vectors_cost is not None else cost_info()
    rc = lib().brille_oracle_sort_pairs(n, pairs.ctypes.data, C.byref(dh.vdesc), C.byref(vc), C.byref(dh.cdesc),
                                        C.byref(cc), row.ctypes.data, col.ctypes.data,
                                        cost.ctypes.data if want_cost else None, int(nthreads))
    return (rc, row, col, cost) if want_cost else (rc, row, col)


def sorted_table(tables, values=None, vectors=None, values_cost=None, vectors_cost=None, nthreads=8):
    """The whole sorted permutation table as {key: permutation}, as
    DualInterpolator::sort leaves it (src/interpolatordual.hpp:398-434)."""
    keys = sort_keys(tables)
    V = np.uint64(np.asarray(tables["vertices"]).shape[0])
    i = (keys // V).astype(np.uint32)
    j = (keys % V).astype(np.uint32)
    fwd = i < j
    pairs = np.stack([i[fwd], j[fwd]], axis=1)
    rc, row, col = sort_pairs(tables, pairs, values, vectors, values_cost, vectors_cost, nthreads)
    return rc, keys, pairs, row, col
