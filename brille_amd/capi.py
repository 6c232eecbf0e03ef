"""ctypes binding of the C ABI declared in include/brille_amd.h.

This is the binding a Python caller (or a test) uses to drive the HIP path
directly through the boundary, without the pybind11 module: structures
mirror the header field by field.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbrille_amd.so")

BRILLE_OK = 0
STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARGUMENT", 2: "NOT_FILLED", 3: "OUTSIDE_BZ", 4: "OUTSIDE_IR", 5: "NULL_NODE",
    6: "NOT_FOUND", 7: "ROTATION", 8: "DEVICE", 9: "NO_DEVICE", 10: "UNSUPPORTED",
}
FP64, FP32 = 0, 1
MEM_DEVICE, MEM_HOST = 0, 1
KEEP_BOTH, KEEP_FP64, KEEP_FP32 = 0, 1, 2
KERNEL_AUTO, KERNEL_PLAIN, KERNEL_TUNED, KERNEL_TILE, KERNEL_BINNED = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)


class GridDesc(C.Structure):
    _fields_ = [
        ("outer_to_xyz", C.c_double * 9), ("outer_metric_recip", C.c_double * 9),
        ("outer_metric_real", C.c_double * 9), ("outer_cross_scale", C.c_double),
        ("transform_needed", C.c_int32), ("to_primitive_6Pt", C.c_int32 * 9),
        ("from_primitive_invPt", C.c_int32 * 9), ("prim_metric_recip", C.c_double * 9),
        ("prim_metric_real", C.c_double * 9), ("prim_cross_scale", C.c_double),
        ("n_faces", C.c_uint32), ("prim_plane_a", _dp), ("prim_plane_b", _dp), ("prim_plane_c", _dp),
        ("prim_normals", _dp), ("prim_taus", _ip), ("prim_tau_lens", _dp),
        ("outer_plane_a", _dp), ("outer_plane_b", _dp), ("outer_plane_c", _dp),
        ("n_wedge", C.c_uint32), ("wedge_normals", _dp), ("no_ir_mirroring", C.c_int32),
        ("has_wedge_seed", C.c_int32), ("wedge_seed", C.c_double * 3),
        ("n_ops", C.c_uint32), ("rotations", _ip), ("inverse_index", _up), ("identity_index", C.c_uint32),
        ("float_tolerance", C.c_double), ("approx_digits", C.c_int32),
        ("n_knots", C.c_uint32 * 3), ("knots", _dp * 3),
        ("n_nodes", C.c_uint64), ("node_type", _u8p), ("node_index", _up),
        ("n_cubes", C.c_uint64), ("cube_vertices", _up),
        ("n_polys", C.c_uint64), ("poly_tet_offset", _up),
        ("n_tets", C.c_uint64), ("tet_vertices", _up), ("tet_circumsphere", _dp), ("tet_volume", _dp),
        ("n_vertices", C.c_uint64), ("vertices", _dp),
    ]


class DataDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_uint64), ("branches", C.c_uint32), ("elements", C.c_uint32 * 3),
                ("rotates_like", C.c_int32), ("length_unit", C.c_int32), ("is_complex", C.c_int32),
                ("data", _dp)]


class GammaDesc(C.Structure):
    _fields_ = [("n_atoms", C.c_uint32), ("n_ops", C.c_uint32), ("F0", _up), ("vector_index", _up),
                ("n_vectors", C.c_uint32), ("vectors", _dp), ("cart_rotations", _dp)]


class PermDesc(C.Structure):
    _fields_ = [("n_keys", C.c_uint64), ("keys", _u64p), ("perm_index", _up), ("n_perms", C.c_uint32),
                ("perms", _up)]


class FillOpts(C.Structure):
    _fields_ = [("memory", C.c_int32), ("data_type", C.c_int32), ("keep", C.c_int32)]


class CostInfo(C.Structure):
    _fields_ = [("scalar_cost", C.c_int32), ("vector_cost", C.c_int32), ("weights", C.c_double * 3)]


def cost_info(scalar_cost=0, vector_cost=0, weights=(1.0, 1.0, 1.0)):
    return CostInfo(int(scalar_cost), int(vector_cost), (C.c_double * 3)(*[float(w) for w in weights]))


class RunOpts(C.Structure):
    _fields_ = [("no_move", C.c_int32), ("precision", C.c_int32), ("use_proposal", C.c_int32),
                ("memory", C.c_int32), ("kernel", C.c_int32), ("check", C.c_int32)]


class AuxOut(C.Structure):
    _fields_ = [("ir_q", C.c_void_p), ("tau", C.c_void_p), ("rot", C.c_void_p), ("invrot", C.c_void_p),
                ("node", C.c_void_p), ("tet", C.c_void_p), ("n_iw", C.c_void_p), ("iw_index", C.c_void_p),
                ("iw_weight", C.c_void_p), ("status", C.c_void_p)]


class TrainOpts(C.Structure):
    _fields_ = [("n_train", C.c_uint64), ("seed", C.c_uint64), ("epochs", C.c_uint32), ("batch", C.c_uint32),
                ("learning_rate", C.c_float)]


class TrainReport(C.Structure):
    _fields_ = [("n_pairs", C.c_uint64), ("final_loss", C.c_float), ("train_hit_rate", C.c_float),
                ("n_classes", C.c_uint32), ("ms", C.c_float)]


class KernelTimes(C.Structure):
    _fields_ = [("n_kernels", C.c_uint32), ("names", (C.c_char * 32) * 4), ("ms", C.c_float * 4),
                ("total_ms", C.c_float)]


# every symbol include/brille_amd.h declares, with its argument types
SIGNATURES = {
    "brille_abi_version": (C.c_int, []),
    "brille_device_count": (C.c_int, []),
    "brille_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "brille_grid_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(C.c_void_p)]),
    "brille_grid_fill": (C.c_int, [C.c_void_p, C.POINTER(DataDesc), C.POINTER(DataDesc), C.POINTER(GammaDesc),
                                   C.POINTER(PermDesc)]),
    "brille_grid_destroy": (C.c_int, [C.c_void_p]),
    "brille_grid_fill_ex": (C.c_int, [C.c_void_p, C.POINTER(DataDesc), C.POINTER(DataDesc), C.POINTER(GammaDesc),
                                      C.POINTER(PermDesc), C.POINTER(FillOpts)]),
    "brille_grid_sort": (C.c_int, [C.c_void_p, C.POINTER(CostInfo), C.POINTER(CostInfo)]),
    "brille_grid_permutations": (C.c_int, [C.c_void_p, _u64p, _up, _u64p, _up, _up]),
    "brille_bytes_per_point": (C.c_uint64, [C.c_void_p, C.c_int32]),
    "brille_ir_interpolate_at": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(RunOpts), C.c_void_p,
                                           C.c_void_p, C.POINTER(AuxOut), C.c_void_p]),
    "brille_interpolate_at": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(RunOpts), C.c_void_p,
                                        C.c_void_p, C.POINTER(AuxOut), C.c_void_p]),
    "brille_ir_interpolate_at_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(RunOpts),
                                                   C.c_void_p, C.c_void_p, C.POINTER(AuxOut), C.c_void_p]),
    "brille_ir_moveinto": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "brille_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_void_p)]),
    "brille_batch_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOpts),
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p]),
    "brille_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOpts),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p]),
    "brille_batch_schedule": (C.c_int, [C.c_void_p, _up, _up, _up, _up, _u64p]),
    "brille_batch_set_schedule": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]),
    "brille_batch_destroy": (C.c_int, [C.c_void_p]),
    "brille_proposal_train": (C.c_int, [C.c_void_p, C.POINTER(TrainOpts), C.POINTER(TrainReport)]),
    "brille_proposal_stats": (C.c_int, [C.c_void_p, _u64p, _u64p]),
    "brille_proposal_loss_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), _ip, C.c_uint32,
                                            C.POINTER(C.c_float), C.POINTER(C.c_float), _up]),
    "brille_proposal_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), _up]),
    "brille_proposal_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint32]),
    "brille_proposal_predict": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), _ip, C.c_uint32, _ip]),
    "brille_time_ir_interpolate_at": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(RunOpts), C.c_void_p,
                                                C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(KernelTimes)]),
    "brille_time_batch_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOpts),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_uint32, C.c_void_p,
                                        C.POINTER(KernelTimes)]),
    "brille_device_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "brille_device_free": (C.c_int, [C.c_void_p]),
    "brille_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "brille_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "brille_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "brille_stream_destroy": (C.c_int, [C.c_void_p]),
    "brille_stream_sync": (C.c_int, [C.c_void_p]),
    "brille_device_sync": (C.c_int, []),
}

_lib = None


def lib():
    """Load libbrille_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python brille_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class BrilleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code
        self.status = STATUS_NAMES.get(code, str(code))


def check(rc):
    if rc != BRILLE_OK:
        buf = C.create_string_buffer(2048)
        lib().brille_last_error(buf, 2048)
        raise BrilleError(rc, buf.value.decode())
    return rc


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class TableHolder:
    """Keeps the numpy arrays a GridDesc points into alive."""

    def __init__(self, tables):
        self.t = {}
        self.desc = GridDesc()
        d = self.desc
        t = tables

        def keep(name, arr, dtype):
            a = np.ascontiguousarray(arr, dtype=dtype)
            self.t[name] = a
            return a

        for name in ("outer_to_xyz", "outer_metric_recip", "outer_metric_real", "prim_metric_recip", "prim_metric_real"):
            getattr(d, name)[:] = [float(x) for x in np.asarray(t[name]).ravel()]
        d.outer_cross_scale = float(t["outer_cross_scale"])
        d.prim_cross_scale = float(t["prim_cross_scale"])
        d.transform_needed = int(t["transform_needed"])
        d.to_primitive_6Pt[:] = [int(x) for x in np.asarray(t["to_primitive_6Pt"]).ravel()]
        d.from_primitive_invPt[:] = [int(x) for x in np.asarray(t["from_primitive_invPt"]).ravel()]
        d.n_faces = int(np.asarray(t["prim_plane_a"]).shape[0])
        for name in ("prim_plane_a", "prim_plane_b", "prim_plane_c", "prim_normals", "prim_tau_lens",
                     "outer_plane_a", "outer_plane_b", "outer_plane_c"):
            setattr(d, name, _ptr(keep(name, t[name], np.float64), C.c_double))
        d.prim_taus = _ptr(keep("prim_taus", t["prim_taus"], np.int32), C.c_int32)
        w = keep("wedge_normals", np.asarray(t["wedge_normals"]).reshape(-1, 3), np.float64)
        d.n_wedge = w.shape[0]
        d.wedge_normals = _ptr(w, C.c_double) if w.shape[0] else None
        d.no_ir_mirroring = int(t["no_ir_mirroring"])
        d.has_wedge_seed = int(t.get("has_wedge_seed", 0))
        d.wedge_seed[:] = [float(x) for x in np.asarray(t.get("wedge_seed", np.zeros(3))).ravel()]
        r = keep("rotations", np.asarray(t["rotations"]).reshape(-1, 9), np.int32)
        d.n_ops = r.shape[0]
        d.rotations = _ptr(r, C.c_int32)
        d.inverse_index = _ptr(keep("inverse_index", t["inverse_index"], np.uint32), C.c_uint32)
        d.identity_index = int(t["identity_index"])
        d.float_tolerance = float(t["float_tolerance"])
        d.approx_digits = int(t["approx_digits"])
        knots = t.get("knots", [np.zeros(0)] * 3)
        for i in range(3):
            k = keep(f"knots{i}", knots[i], np.float64)
            d.n_knots[i] = k.shape[0]
            d.knots[i] = _ptr(k, C.c_double) if k.shape[0] else None
        if "node_type" in t:
            nt = keep("node_type", t["node_type"], np.uint8)
            d.n_nodes = nt.shape[0]
            d.node_type = _ptr(nt, C.c_uint8)
            d.node_index = _ptr(keep("node_index", t["node_index"], np.uint32), C.c_uint32)
            cv = keep("cube_vertices", np.asarray(t["cube_vertices"]).reshape(-1, 8), np.uint32)
            d.n_cubes = cv.shape[0]
            d.cube_vertices = _ptr(cv, C.c_uint32)
            po = keep("poly_tet_offset", t["poly_tet_offset"], np.uint32)
            d.n_polys = max(po.shape[0] - 1, 0)
            d.poly_tet_offset = _ptr(po, C.c_uint32)
            tv = keep("tet_vertices", np.asarray(t["tet_vertices"]).reshape(-1, 4), np.uint32)
            d.n_tets = tv.shape[0]
            d.tet_vertices = _ptr(tv, C.c_uint32)
            d.tet_circumsphere = _ptr(keep("tet_circumsphere", np.asarray(t["tet_circumsphere"]).reshape(-1, 4), np.float64), C.c_double)
            d.tet_volume = _ptr(keep("tet_volume", t["tet_volume"], np.float64), C.c_double)
            vx = keep("vertices", np.asarray(t["vertices"]).reshape(-1, 3), np.float64)
            d.n_vertices = vx.shape[0]
            d.vertices = _ptr(vx, C.c_double)


class DataHolder:
    """DataDesc/GammaDesc/PermDesc for a filled grid's payload."""

    def __init__(self, tables, values=None, vectors=None, perm=None):
        t = tables
        vals = np.ascontiguousarray(t["values"] if values is None else values, dtype=np.float64)
        vecs = np.ascontiguousarray(t["vectors"] if vectors is None else vectors, dtype=np.complex128)
        self.vals, self.vecs = vals, vecs
        self.vdesc = DataDesc()
        self.cdesc = DataDesc()
        B = int(t["branches"])
        for desc, arr, el, rot, lu, cplx in (
                (self.vdesc, vals, t["values_elements"], t["values_rotates_like"], t["values_length_unit"], 0),
                (self.cdesc, vecs, t["vectors_elements"], t["vectors_rotates_like"], t["vectors_length_unit"], 1)):
            desc.n_vertices = arr.shape[0]
            desc.branches = B
            desc.elements[:] = [int(x) for x in el]
            desc.rotates_like = int(rot)
            desc.length_unit = int(lu)
            desc.is_complex = cplx
            desc.data = arr.ctypes.data_as(_dp)
        self.gdesc = None
        if "gamma_F0" in t:
            self.F0 = np.ascontiguousarray(t["gamma_F0"], dtype=np.uint32)
            self.vi = np.ascontiguousarray(t["gamma_vector_index"], dtype=np.uint32)
            self.gv = np.ascontiguousarray(t["gamma_vectors"], dtype=np.float64)
            self.cr = np.ascontiguousarray(t["gamma_cart_rotations"], dtype=np.float64)
            g = GammaDesc()
            g.n_atoms, g.n_ops = self.F0.shape
            g.F0 = _ptr(self.F0, C.c_uint32)
            g.vector_index = _ptr(self.vi, C.c_uint32)
            g.n_vectors = self.gv.shape[0]
            g.vectors = _ptr(self.gv, C.c_double)
            g.cart_rotations = _ptr(self.cr, C.c_double)
            self.gdesc = g
        self.pdesc = None
        if perm is not None:
            keys, index, perms = perm
            self.pk = np.ascontiguousarray(keys, dtype=np.uint64)
            self.pi = np.ascontiguousarray(index, dtype=np.uint32)
            self.pp = np.ascontiguousarray(perms, dtype=np.uint32)
            p = PermDesc()
            p.n_keys = self.pk.shape[0]
            p.keys = _ptr(self.pk, C.c_uint64)
            p.perm_index = _ptr(self.pi, C.c_uint32)
            p.n_perms = self.pp.shape[0]
            p.perms = _ptr(self.pp, C.c_uint32)
            self.pdesc = p


class DeviceArray:
    """A device allocation made through the C ABI."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().brille_device_alloc(max(self.nbytes, 1), C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().brille_memcpy_h2d(d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().brille_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def rows_to_host(self, r0, r1):
        """Rows [r0, r1) of the leading axis, copied to the host."""
        row = self.nbytes // max(self.shape[0], 1)
        out = np.empty((r1 - r0,) + self.shape[1:], self.dtype)
        if out.nbytes:
            check(lib().brille_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr.value + r0 * row),
                                          out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().brille_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Grid:
    """A grid handle created from a tables dict (BZTrellisQdc.tables())."""

    def __init__(self, tables, values=None, vectors=None, perm=None):
        self.tables = TableHolder(tables)
        h = C.c_void_p()
        check(lib().brille_grid_create(C.byref(self.tables.desc), C.byref(h)))
        self.h = h
        self.data = None
        if "values" in tables or values is not None:
            self.fill(tables, values, vectors, perm)

    def fill(self, tables, values=None, vectors=None, perm=None):
        self.data = DataHolder(tables, values, vectors, perm)
        d = self.data
        check(lib().brille_grid_fill(self.h, C.byref(d.vdesc), C.byref(d.cdesc),
                                     C.byref(d.gdesc) if d.gdesc is not None else None,
                                     C.byref(d.pdesc) if d.pdesc is not None else None))

    def fill_ex(self, tables, values_ptr, vectors_ptr, memory, data_type, keep=KEEP_BOTH, perm=None):
        """brille_grid_fill_ex: the payload's element layout and GammaTable come from `tables`
        (its own values/vectors arrays are not read); values_ptr/vectors_ptr point to the
        payload, float/complex64 when data_type is FP32, in host or device memory."""
        V = int(np.asarray(tables["vertices"]).shape[0])
        B = int(tables["branches"])
        sv = int(np.sum(tables["values_elements"]))
        sc = int(np.sum(tables["vectors_elements"]))
        shape_v = (V, B, sv)
        shape_c = (V, B, sc)
        stub = dict(tables)
        stub["values"] = np.zeros((1,) + shape_v[1:])
        stub["vectors"] = np.zeros((1,) + shape_c[1:], complex)
        d = DataHolder(stub, perm=perm)
        for desc, ptr in ((d.vdesc, values_ptr), (d.cdesc, vectors_ptr)):
            desc.n_vertices = V
            desc.data = C.cast(C.c_void_p(int(ptr)), _dp)
        self.data = d
        self.data.vals = np.empty(shape_v, np.float32)[:0]  # shapes for out_shapes
        self.data.vals_shape, self.data.vecs_shape = shape_v, shape_c
        o = FillOpts(memory, data_type, keep)
        check(lib().brille_grid_fill_ex(self.h, C.byref(d.vdesc), C.byref(d.cdesc),
                                        C.byref(d.gdesc) if d.gdesc is not None else None,
                                        C.byref(d.pdesc) if d.pdesc is not None else None, C.byref(o)))

    def sort(self, values_cost=None, vectors_cost=None):
        """brille_grid_sort: mode permutations of every connected vertex pair."""
        check(lib().brille_grid_sort(self.h, C.byref(values_cost) if values_cost is not None else None,
                                     C.byref(vectors_cost) if vectors_cost is not None else None))

    def permutations(self):
        """(keys, perm_index, perms) of the grid's permutation table."""
        nk, npm = C.c_uint64(), C.c_uint32()
        L = lib()
        check(L.brille_grid_permutations(self.h, C.byref(nk), C.byref(npm), None, None, None))
        keys = np.zeros(nk.value, np.uint64)
        index = np.zeros(nk.value, np.uint32)
        perms = np.zeros((npm.value, self.branches), np.uint32)
        check(L.brille_grid_permutations(self.h, C.byref(nk), C.byref(npm), keys.ctypes.data_as(_u64p),
                                         index.ctypes.data_as(_up), perms.ctypes.data_as(_up)))
        return keys, index, perms

    @property
    def branches(self):
        return self.data.vdesc.branches

    def out_shapes(self, n):
        if self.data is None:  # unfilled: the library reports it before touching the outputs
            return (n, 0), (n, 0)
        vs = (n,) + tuple(getattr(self.data, "vals_shape", self.data.vals.shape)[1:])
        cs = (n,) + tuple(getattr(self.data, "vecs_shape", self.data.vecs.shape)[1:])
        return vs, cs

    def interpolate_at(self, Q, precision=FP64, kernel=KERNEL_AUTO, no_move=False, aux=False):
        """The first-zone variant (brille_interpolate_at); same returns as ir_interpolate_at."""
        return self.ir_interpolate_at(Q, precision, kernel, no_move, aux, False, _entry="brille_interpolate_at")

    def ir_interpolate_at(self, Q, precision=FP64, kernel=KERNEL_AUTO, no_move=False, aux=False, use_proposal=False,
                          _entry="brille_ir_interpolate_at"):
        """Host-memory call; returns (values, vectors[, aux dict]).  aux: True for
        every per-point output, or a dict naming outputs to leave out."""
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 3)
        n = Q.shape[0]
        vs, cs = self.out_shapes(n)
        vals = np.zeros(vs, np.float32 if precision == FP32 else np.float64)
        vecs = np.zeros(cs, np.complex64 if precision == FP32 else np.complex128)
        o = RunOpts(int(no_move), precision, int(use_proposal), MEM_HOST, kernel, 1)
        a = None
        bufs = {}
        if aux:
            bufs = dict(ir_q=np.zeros((n, 3)), tau=np.zeros((n, 3), np.int32), rot=np.zeros(n, np.uint32),
                        invrot=np.zeros(n, np.uint32), node=np.zeros(n, np.uint32), tet=np.zeros(n, np.uint32),
                        n_iw=np.zeros(n, np.uint8), iw_index=np.zeros((n, 8), np.uint32),
                        iw_weight=np.zeros((n, 8)), status=np.zeros(n, np.uint8))
            if isinstance(aux, dict):  # {name: False} leaves an output out (the binned pipeline has no tau)
                bufs = {k: v for k, v in bufs.items() if aux.get(k, True)}
            a = AuxOut(**{k: v.ctypes.data_as(C.c_void_p) for k, v in bufs.items()})
        check(getattr(lib(), _entry)(self.h, Q.ctypes.data_as(C.c_void_p), n, C.byref(o),
                                     vals.ctypes.data_as(C.c_void_p), vecs.ctypes.data_as(C.c_void_p),
                                     C.byref(a) if a is not None else None, None))
        if aux:
            return vals, vecs, bufs
        return vals, vecs

    def run_device(self, Q, precision=FP64, kernel=KERNEL_AUTO, no_move=False, check=True, use_proposal=False,
                   aux=("rot", "invrot", "node", "tet", "n_iw", "iw_index", "iw_weight", "status"),
                   _entry="brille_ir_interpolate_at"):
        """Device-memory call: Q and every output live in HBM for the call.  With
        check=False no error is reduced and the per-point status is returned instead.
        Returns (rc, values, vectors, aux dict) on the host."""
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 3)
        n = Q.shape[0]
        vs, cs = self.out_shapes(n)
        dQ = DeviceArray.from_host(Q)
        dv = DeviceArray(vs, np.float32 if precision == FP32 else np.float64)
        dc = DeviceArray(cs, np.complex64 if precision == FP32 else np.complex128)
        shapes = dict(ir_q=((n, 3), np.float64), tau=((n, 3), np.int32), rot=((n,), np.uint32),
                      invrot=((n,), np.uint32), node=((n,), np.uint32), tet=((n,), np.uint32),
                      n_iw=((n,), np.uint8), iw_index=((n, 8), np.uint32), iw_weight=((n, 8), np.float64),
                      status=((n,), np.uint8))
        dev = {k: DeviceArray(*shapes[k]) for k in aux}
        a = AuxOut(**{k: v.ptr for k, v in dev.items()}) if dev else None
        o = RunOpts(int(no_move), precision, int(use_proposal), MEM_DEVICE, kernel, int(check))
        rc = getattr(lib(), _entry)(self.h, dQ.ptr, n, C.byref(o), dv.ptr, dc.ptr,
                                    C.byref(a) if a is not None else None, None)
        lib().brille_device_sync()
        return rc, dv.to_host(), dc.to_host(), {k: v.to_host() for k, v in dev.items()}

    def ir_moveinto(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 3)
        n = Q.shape[0]
        q = np.zeros((n, 3))
        tau = np.zeros((n, 3), np.int32)
        r = np.zeros(n, np.uint32)
        ir = np.zeros(n, np.uint32)
        st = np.zeros(n, np.uint8)
        rc = lib().brille_ir_moveinto(self.h, Q.ctypes.data_as(C.c_void_p), n, MEM_HOST,
                                      q.ctypes.data_as(C.c_void_p), tau.ctypes.data_as(C.c_void_p),
                                      r.ctypes.data_as(C.c_void_p), ir.ctypes.data_as(C.c_void_p),
                                      st.ctypes.data_as(C.c_void_p), None)
        return rc, q, tau, r, ir, st

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().brille_grid_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Batch:
    """Many grid handles run in one persistent launch (brille_batch_*), on
    device buffers: Q of all grids back to back, one output pair per grid."""

    def __init__(self, grids):
        self.grids = list(grids)  # the handles must outlive the batch
        arr = (C.c_void_p * len(self.grids))(*[g.h for g in self.grids])
        h = C.c_void_p()
        check(lib().brille_batch_create(arr, len(self.grids), C.byref(h)))
        self.h = h

    def outputs(self, offsets, precision):
        """Device output buffers for every grid."""
        vd = np.float32 if precision == FP32 else np.float64
        cd = np.complex64 if precision == FP32 else np.complex128
        vals, vecs = [], []
        for g, grid in enumerate(self.grids):
            n = int(offsets[g + 1] - offsets[g])
            vs, cs = grid.out_shapes(n)
            vals.append(DeviceArray(vs, vd))
            vecs.append(DeviceArray(cs, cd))
        return vals, vecs

    def _args(self, offsets, vals, vecs):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        vp = (C.c_void_p * len(vals))(*[v.ptr for v in vals])
        cp = (C.c_void_p * len(vecs))(*[c.ptr for c in vecs])
        return off, vp, cp

    def run(self, dQ, offsets, vals, vecs, precision=FP64, kernel=KERNEL_AUTO, check_errors=True, autotune=False):
        off, vp, cp = self._args(offsets, vals, vecs)
        o = RunOpts(0, precision, 0, MEM_DEVICE, kernel, int(check_errors))
        f = lib().brille_batch_autotune if autotune else lib().brille_batch_run
        return f(self.h, dQ.ptr, off.ctypes.data_as(_u64p), C.byref(o), vp, cp, None)

    def time(self, dQ, offsets, vals, vecs, precision=FP32, kernel=KERNEL_AUTO, iters=3):
        off, vp, cp = self._args(offsets, vals, vecs)
        o = RunOpts(0, precision, 0, MEM_DEVICE, kernel, 0)
        t = KernelTimes()
        check(lib().brille_time_batch_run(self.h, dQ.ptr, off.ctypes.data_as(_u64p), C.byref(o), vp, cp, iters,
                                          None, C.byref(t)))
        return t.total_ms

    def set_schedule(self, tile_points=256, blocks_per_cu=4, tiles_per_grab=1, order=1, binned_min=None):
        check(lib().brille_batch_set_schedule(self.h, tile_points, blocks_per_cu, tiles_per_grab, order,
                                              2 ** 64 - 1 if binned_min is None else int(binned_min)))

    def schedule(self):
        v = [C.c_uint32() for _ in range(4)]
        bmin = C.c_uint64()
        check(lib().brille_batch_schedule(self.h, *[C.byref(x) for x in v], C.byref(bmin)))
        return dict(tile_points=v[0].value, blocks=v[1].value, tiles_per_grab=v[2].value, order=v[3].value,
                    binned_min=None if bmin.value == 2 ** 64 - 1 else int(bmin.value))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().brille_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass
