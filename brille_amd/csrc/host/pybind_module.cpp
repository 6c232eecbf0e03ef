// pybind11 module `_brille`: the Python surface of the path, with the
// reference's class and argument names (wrap/_trellis.hpp:29-103,
// wrap/_common_grid.hpp:20-341, wrap/_bz.cpp:156-497, wrap/_lattice.cpp:75-150,
// wrap/_approx.cpp:7-30).  Hot-path calls go through the C ABI in
// include/brille_amd.h; nothing here computes on the CPU in place of a kernel.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstring>

#include "grid.hpp"

namespace py = pybind11;
using namespace pybind11::literals;
using namespace brille_amd;

namespace {

using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using iarr = py::array_t<int, py::array::c_style | py::array::forcecast>;

V3 to_v3(const darr& a) {
  if (a.size() != 3) throw std::runtime_error("expected three values");
  const double* p = a.data();
  return {p[0], p[1], p[2]};
}

py::array_t<double> v3list(const std::vector<V3>& v) {
  py::array_t<double> out({static_cast<py::ssize_t>(v.size()), static_cast<py::ssize_t>(3)});
  auto r = out.mutable_unchecked<2>();
  for (size_t i = 0; i < v.size(); ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = v[i][j];
  return out;
}

py::array_t<double> mat33(const M3& m) {
  py::array_t<double> out({3, 3});
  std::memcpy(out.mutable_data(), m.data(), 9 * sizeof(double));
  return out;
}

template <class T>
py::array_t<T> vec_array(const std::vector<T>& v, std::vector<py::ssize_t> shape) {
  py::array_t<T> out(shape);
  if (!v.empty()) std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(T));
  return out;
}

Symmetry symmetry_from_string(const std::string& s, Bravais* b) { return spacegroup_from_string(s, b); }

Basis basis_or_empty(const py::object& o) {
  if (o.is_none()) return Basis();
  return o.cast<Basis>();
}

// BZTrellisQdc plus the Python objects handed to fill (held by reference
// as the reference does, wrap/tests/test_4_interpolation.py:265-277)
struct PyGrid {
  std::shared_ptr<BZTrellisQdc> g;
  py::object values_ref, vectors_ref;
};

struct PyMesh : PyGrid {};
struct PyNest : PyGrid {};

struct PyZone {
  BrillouinZone bz;
  std::shared_ptr<ZoneTables> tables;
  std::shared_ptr<DeviceGrid> dev;
  brille_grid* handle() {
    if (!dev) {
      tables = std::make_shared<ZoneTables>();
      dev = zone_handle(bz, *tables);
    }
    return dev->h;
  }
};

Payload payload_from(const py::array& arr, const iarr& el, bool complex_data) {
  std::vector<size_t> shape;
  for (py::ssize_t i = 0; i < arr.ndim(); ++i) shape.push_back(static_cast<size_t>(arr.shape(i)));
  std::vector<int> e(el.data(), el.data() + el.size());
  if (el.ndim() != 1) throw std::runtime_error("elements must be a 1-D array");
  Payload p;
  parse_payload(p, shape, e, complex_data);
  size_t total = 1;
  for (auto s : shape) total *= s;
  if (complex_data) {
    auto c = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>::ensure(arr);
    if (!c) throw std::runtime_error("vectors must be convertible to complex128");
    p.data.resize(2 * total);
    std::memcpy(p.data.data(), c.data(), total * sizeof(std::complex<double>));
  } else {
    auto r = darr::ensure(arr);
    if (!r) throw std::runtime_error("values must be convertible to float64");
    p.data.assign(r.data(), r.data() + total);
  }
  return p;
}

// the cost functions and weights of a weighted fill  wrap/_interpolator.hpp:96-105
void set_cost(Payload& p, const iarr& el, const darr& w) {
  if (w.ndim() != 1) throw std::runtime_error("weights must be a 1-D array");
  if (el.size() > 5) p.cost.scalar_cost = el.data()[5];
  if (el.size() > 6) p.cost.vector_cost = el.data()[6];
  for (py::ssize_t i = 0; i < w.size() && i < 3; ++i) p.cost.weights[i] = w.data()[i];
}

py::tuple grid_interpolate(PyGrid& pg, const darr& Q, bool useparallel, int threads, bool no_move,
                           const std::string& precision, const std::string& kernel, bool first_zone) {
  (void)useparallel;
  (void)threads;
  if (Q.ndim() != 2 || Q.shape(1) != 3) throw std::runtime_error("Interpolation requires one or more 3-vectors");
  BZTrellisQdc& g = *pg.g;
  if (!g.filled) throw std::runtime_error("The interpolation data must be filled before interpolating.");
  const int prec = precision == "fp32" ? BRILLE_FP32 : BRILLE_FP64;
  if (precision != "fp32" && precision != "fp64") throw std::runtime_error("precision must be 'fp64' or 'fp32'");
  int kern = BRILLE_KERNEL_AUTO;
  if (kernel == "plain") kern = BRILLE_KERNEL_PLAIN;
  else if (kernel == "tuned") kern = BRILLE_KERNEL_TUNED;
  else if (kernel == "tile") kern = BRILLE_KERNEL_TILE;
  else if (kernel == "binned") kern = BRILLE_KERNEL_BINNED;
  const size_t n = static_cast<size_t>(Q.shape(0));
  std::vector<py::ssize_t> vs{static_cast<py::ssize_t>(n)}, cs{static_cast<py::ssize_t>(n)};
  for (size_t i = 1; i < g.values.shape.size(); ++i) vs.push_back(static_cast<py::ssize_t>(g.values.shape[i]));
  for (size_t i = 1; i < g.vectors.shape.size(); ++i) cs.push_back(static_cast<py::ssize_t>(g.vectors.shape[i]));
  py::array vals, vecs;
  if (prec == BRILLE_FP32) {
    vals = py::array_t<float>(vs);
    vecs = py::array_t<std::complex<float>>(cs);
  } else {
    vals = py::array_t<double>(vs);
    vecs = py::array_t<std::complex<double>>(cs);
  }
  {
    py::gil_scoped_release release;
    g.ir_interpolate_at(Q.data(), n, no_move, prec, kern, vals.mutable_data(), vecs.mutable_data(), nullptr,
                        first_zone);
  }
  return py::make_tuple(vals, vecs);
}

py::dict grid_tables(const BZTrellisQdc& g) {
  brille_grid_desc d = g.describe();
  py::dict t;
  auto arr = [](const double* p, size_t n, std::vector<py::ssize_t> shape) {
    py::array_t<double> a(shape);
    if (n) std::memcpy(a.mutable_data(), p, n * sizeof(double));
    return a;
  };
  t["outer_to_xyz"] = arr(d.outer_to_xyz, 9, {9});
  t["outer_metric_recip"] = arr(d.outer_metric_recip, 9, {9});
  t["outer_metric_real"] = arr(d.outer_metric_real, 9, {9});
  t["outer_cross_scale"] = d.outer_cross_scale;
  t["transform_needed"] = d.transform_needed;
  t["to_primitive_6Pt"] = py::array_t<int32_t>({9}, d.to_primitive_6Pt);
  t["from_primitive_invPt"] = py::array_t<int32_t>({9}, d.from_primitive_invPt);
  t["prim_metric_recip"] = arr(d.prim_metric_recip, 9, {9});
  t["prim_metric_real"] = arr(d.prim_metric_real, 9, {9});
  t["prim_cross_scale"] = d.prim_cross_scale;
  const py::ssize_t F = d.n_faces;
  t["prim_plane_a"] = arr(d.prim_plane_a, 3 * F, {F, 3});
  t["prim_plane_b"] = arr(d.prim_plane_b, 3 * F, {F, 3});
  t["prim_plane_c"] = arr(d.prim_plane_c, 3 * F, {F, 3});
  t["prim_normals"] = arr(d.prim_normals, 3 * F, {F, 3});
  t["prim_taus"] = py::array_t<int32_t>({F, (py::ssize_t)3}, d.prim_taus);
  t["prim_tau_lens"] = arr(d.prim_tau_lens, F, {F});
  t["outer_plane_a"] = arr(d.outer_plane_a, 3 * F, {F, 3});
  t["outer_plane_b"] = arr(d.outer_plane_b, 3 * F, {F, 3});
  t["outer_plane_c"] = arr(d.outer_plane_c, 3 * F, {F, 3});
  const py::ssize_t W = d.n_wedge;
  t["wedge_normals"] = arr(d.wedge_normals, 3 * W, {W, 3});
  t["no_ir_mirroring"] = d.no_ir_mirroring;
  t["has_wedge_seed"] = d.has_wedge_seed;
  t["wedge_seed"] = arr(d.wedge_seed, 3, {3});
  const py::ssize_t G = d.n_ops;
  t["rotations"] = py::array_t<int32_t>({G, (py::ssize_t)9}, d.rotations);
  t["inverse_index"] = py::array_t<uint32_t>({G}, d.inverse_index);
  t["identity_index"] = d.identity_index;
  t["float_tolerance"] = d.float_tolerance;
  t["approx_digits"] = d.approx_digits;
  py::list knots;
  for (int i = 0; i < 3; ++i) knots.append(arr(d.knots[i], d.n_knots[i], {(py::ssize_t)d.n_knots[i]}));
  t["knots"] = knots;
  const auto& T = g.trellis;
  t["node_type"] = vec_array(T.node_type, {(py::ssize_t)T.node_type.size()});
  t["node_index"] = vec_array(T.node_index, {(py::ssize_t)T.node_index.size()});
  t["cube_vertices"] = vec_array(T.cube_vertices, {(py::ssize_t)T.n_cubes(), 8});
  t["poly_tet_offset"] = vec_array(T.poly_offset, {(py::ssize_t)T.poly_offset.size()});
  t["poly_node"] = vec_array(T.poly_node, {(py::ssize_t)T.poly_node.size()});
  t["tet_vertices"] = vec_array(T.tet_vertices, {(py::ssize_t)T.n_tets(), 4});
  t["tet_circumsphere"] = vec_array(T.tet_circum, {(py::ssize_t)T.n_tets(), 4});
  t["tet_volume"] = vec_array(T.tet_volume, {(py::ssize_t)T.n_tets()});
  t["vertices"] = vec_array(T.vertices, {(py::ssize_t)T.n_vertices(), 3});
  if (g.filled) {
    std::vector<py::ssize_t> vs, cs;
    for (auto s : g.values.shape) vs.push_back((py::ssize_t)s);
    for (auto s : g.vectors.shape) cs.push_back((py::ssize_t)s);
    t["values"] = vec_array(g.values.data, vs);
    py::array_t<std::complex<double>> c(cs);
    std::memcpy(c.mutable_data(), g.vectors.data.data(), g.vectors.data.size() * sizeof(double));
    t["vectors"] = c;
    t["values_elements"] = std::vector<uint32_t>{g.values.elements[0], g.values.elements[1], g.values.elements[2]};
    t["vectors_elements"] = std::vector<uint32_t>{g.vectors.elements[0], g.vectors.elements[1], g.vectors.elements[2]};
    t["values_rotates_like"] = g.values.rot;
    t["values_length_unit"] = g.values.lu;
    t["vectors_rotates_like"] = g.vectors.rot;
    t["vectors_length_unit"] = g.vectors.lu;
    t["branches"] = g.values.branches;
    if (g.vectors.rot == BRILLE_ROT_GAMMA) {
      t["gamma_F0"] = vec_array(g.gamma.F0, {(py::ssize_t)g.gamma.n_atoms, (py::ssize_t)g.gamma.n_ops});
      t["gamma_vector_index"] = vec_array(g.gamma.vindex, {(py::ssize_t)g.gamma.n_atoms, (py::ssize_t)g.gamma.n_ops});
      t["gamma_vectors"] = vec_array(g.gamma.vectors, {(py::ssize_t)(g.gamma.vectors.size() / 3), 3});
      t["gamma_cart_rotations"] = vec_array(g.gamma.cart_rotations, {(py::ssize_t)g.gamma.n_ops, 9});
    }
  }
  return t;
}

py::object make_poly(const ConvexPolyhedron& P) {
  py::dict d;
  return py::cast(P);
}

}  // namespace

PYBIND11_MODULE(_brille, m) {
  m.doc() = "MI355X implementation of brille's BZTrellisQdc.ir_interpolate_at path";

  m.attr("RotatesLike") = py::dict("vector"_a = 0, "pseudovector"_a = 1, "Gamma"_a = 2);
  m.attr("LengthUnit") = py::dict("none"_a = 0, "angstrom"_a = 1, "inverse_angstrom"_a = 2, "real_lattice"_a = 3,
                                  "reciprocal_lattice"_a = 4);
  m.attr("NodeType") = py::dict("assumed_null"_a = 0, "found_null"_a = 1, "null"_a = 2, "cube"_a = 3, "poly"_a = 4);

  py::class_<ApproxConfig>(m, "ApproxConfig")
      .def(py::init([](int dig, double dir, double rec) { return ApproxConfig{dig, dir, rec}; }), "digits"_a = 1000,
           "real_space_tolerance"_a = 1e-10, "reciprocal_space_tolerance"_a = 1e-10)
      .def_readwrite("digits", &ApproxConfig::digits)
      .def_readwrite("real_space_tolerance", &ApproxConfig::direct)
      .def_readwrite("reciprocal_space_tolerance", &ApproxConfig::reciprocal);

  py::class_<Basis>(m, "Basis")
      .def(py::init<>())
      .def(py::init([](const darr& pos, const py::array_t<long long, py::array::c_style | py::array::forcecast>& typ) {
             if (pos.ndim() != 2 || pos.shape(1) != 3) throw std::runtime_error("positions must be (n, 3)");
             std::vector<V3> p;
             std::vector<uint32_t> t;
             for (py::ssize_t i = 0; i < pos.shape(0); ++i) p.push_back({pos.at(i, 0), pos.at(i, 1), pos.at(i, 2)});
             for (py::ssize_t i = 0; i < typ.size(); ++i) t.push_back(static_cast<uint32_t>(typ.data()[i]));
             if (t.size() != p.size()) throw std::runtime_error("one type per position required");
             return Basis(p, t);
           }),
           "positions"_a, "types"_a)
      .def_property_readonly("positions", [](const Basis& b) { return v3list(b.positions); })
      .def_property_readonly("types", [](const Basis& b) { return b.types; })
      .def("__len__", &Basis::size);

  py::class_<Symmetry>(m, "Symmetry")
      .def(py::init([](const py::array_t<int, py::array::c_style | py::array::forcecast>& W, const darr& w) {
             if (W.ndim() != 3 || W.shape(1) != 3 || W.shape(2) != 3) throw std::runtime_error("W must be (n,3,3)");
             if (w.ndim() != 2 || w.shape(1) != 3 || w.shape(0) != W.shape(0)) throw std::runtime_error("w must be (n,3)");
             std::vector<IM3> Wm;
             std::vector<V3> wv;
             for (py::ssize_t i = 0; i < W.shape(0); ++i) {
               IM3 r;
               for (int j = 0; j < 9; ++j) r[j] = W.data()[9 * i + j];
               Wm.push_back(r);
               wv.push_back({w.at(i, 0), w.at(i, 1), w.at(i, 2)});
             }
             return Symmetry::from_matrices(Wm, wv);
           }),
           "W"_a, "w"_a)
      .def(py::init([](const std::string& s) { return spacegroup_from_string(s); }), "symbol"_a)
      .def_property_readonly("W", [](const Symmetry& s) {
        py::array_t<int> out({(py::ssize_t)s.size(), (py::ssize_t)3, (py::ssize_t)3});
        for (size_t i = 0; i < s.size(); ++i) std::memcpy(out.mutable_data() + 9 * i, s.ops[i].W.data(), 9 * sizeof(int));
        return out;
      })
      .def_property_readonly("w", [](const Symmetry& s) {
        std::vector<V3> v;
        for (const auto& o : s.ops) v.push_back(o.w);
        return v3list(v);
      })
      .def_property_readonly("size", &Symmetry::size)
      .def("generate", &Symmetry::generate)
      .def("__len__", &Symmetry::size);

  py::class_<PointSymmetry>(m, "PointSymmetry")
      .def_property_readonly("W", [](const PointSymmetry& p) {
        py::array_t<int> out({(py::ssize_t)p.size(), (py::ssize_t)3, (py::ssize_t)3});
        for (size_t i = 0; i < p.size(); ++i) std::memcpy(out.mutable_data() + 9 * i, p.R[i].data(), 9 * sizeof(int));
        return out;
      })
      .def_property_readonly("size", &PointSymmetry::size)
      .def("find_identity_index", &PointSymmetry::find_identity_index)
      .def("get_inverse_index", &PointSymmetry::get_inverse_index)
      .def("has_space_inversion", &PointSymmetry::has_space_inversion)
      .def("__len__", &PointSymmetry::size);

  py::class_<Lattice>(m, "Lattice")
      .def(py::init([](const darr& lengths, const darr& angles, const std::string& sym, const py::object& basis, bool snap,
                       bool real_space) {
             (void)snap;
             Bravais b;
             Symmetry s = symmetry_from_string(sym, &b);
             return Lattice::from_parameters(to_v3(lengths), to_v3(angles), s, s.centring(), basis_or_empty(basis), real_space);
           }),
           "basis_vector_lengths"_a, "basis_vector_angles"_a, "symmetry_information"_a = "P 1", "basis"_a = py::none(),
           py::kw_only(), "snap_to_symmetry"_a = true, "real_space"_a = true)
      .def(py::init([](const darr& lengths, const darr& angles, const Symmetry& sym, const py::object& basis, bool snap,
                       bool real_space) {
             (void)snap;
             return Lattice::from_parameters(to_v3(lengths), to_v3(angles), sym, sym.centring(), basis_or_empty(basis), real_space);
           }),
           "basis_vector_lengths"_a, "basis_vector_angles"_a, "symmetry"_a, "basis"_a = py::none(), py::kw_only(),
           "snap_to_symmetry"_a = true, "real_space"_a = true)
      .def(py::init([](const darr& vectors, const Symmetry& sym, const py::object& basis, bool snap, bool real_space,
                       bool row_vectors) {
             (void)snap;
             if (vectors.size() != 9) throw std::runtime_error("basis_vectors must be 3x3");
             M3 v;
             std::memcpy(v.data(), vectors.data(), 9 * sizeof(double));
             return Lattice::from_vectors(v, row_vectors, sym, sym.centring(), basis_or_empty(basis), real_space);
           }),
           "basis_vectors"_a, "symmetry"_a, "basis"_a = py::none(), py::kw_only(), "snap_to_symmetry"_a = true,
           "real_space"_a = true, "row_vectors"_a = true)
      .def(py::init([](const darr& vectors, const std::string& sym, const py::object& basis, bool snap, bool real_space,
                       bool row_vectors) {
             (void)snap;
             if (vectors.size() != 9) throw std::runtime_error("basis_vectors must be 3x3");
             M3 v;
             std::memcpy(v.data(), vectors.data(), 9 * sizeof(double));
             Symmetry s = symmetry_from_string(sym, nullptr);
             return Lattice::from_vectors(v, row_vectors, s, s.centring(), basis_or_empty(basis), real_space);
           }),
           "basis_vectors"_a, "symmetry_information"_a = "P 1", "basis"_a = py::none(), py::kw_only(),
           "snap_to_symmetry"_a = true, "real_space"_a = true, "row_vectors"_a = true)
      .def_property_readonly("real_vectors", [](const Lattice& l) { return mat33(l.real); })
      .def_property_readonly("reciprocal_vectors", [](const Lattice& l) { return mat33(l.recip); })
      .def_property_readonly("a", [](const Lattice& l) { return l.lengths(true)[0]; })
      .def_property_readonly("b", [](const Lattice& l) { return l.lengths(true)[1]; })
      .def_property_readonly("c", [](const Lattice& l) { return l.lengths(true)[2]; })
      .def_property_readonly("alpha", [](const Lattice& l) { return l.angles(true)[0]; })
      .def_property_readonly("beta", [](const Lattice& l) { return l.angles(true)[1]; })
      .def_property_readonly("gamma", [](const Lattice& l) { return l.angles(true)[2]; })
      .def_property_readonly("volume", &Lattice::volume_real)
      .def_property_readonly("a_star", [](const Lattice& l) { return l.lengths(false)[0]; })
      .def_property_readonly("b_star", [](const Lattice& l) { return l.lengths(false)[1]; })
      .def_property_readonly("c_star", [](const Lattice& l) { return l.lengths(false)[2]; })
      .def_property_readonly("volume_star", &Lattice::volume_recip)
      .def_property_readonly("bravais", [](const Lattice& l) { return static_cast<int>(l.bravais); })
      .def_property_readonly("spacegroup", [](const Lattice& l) { return l.space; })
      .def_property_readonly("pointgroup", [](const Lattice& l) { return l.point; })
      .def_property_readonly("basis", [](const Lattice& l) { return l.basis; })
      .def("get_covariant_metric_tensor", [](const Lattice& l) { return mat33(l.real_metric); })
      .def("get_contravariant_metric_tensor", [](const Lattice& l) { return mat33(l.recip_metric); });

  m.def("Direct", [](const darr& lengths, const darr& angles, const std::string& sym, const py::object& basis) {
    Symmetry s = symmetry_from_string(sym, nullptr);
    return Lattice::from_parameters(to_v3(lengths), to_v3(angles), s, s.centring(), basis_or_empty(basis), true);
  }, "lengths"_a, "angles"_a, "symmetry_information"_a = "P 1", "basis"_a = py::none());
  m.def("Reciprocal", [](const darr& lengths, const darr& angles, const std::string& sym, const py::object& basis) {
    Symmetry s = symmetry_from_string(sym, nullptr);
    return Lattice::from_parameters(to_v3(lengths), to_v3(angles), s, s.centring(), basis_or_empty(basis), false);
  }, "lengths"_a, "angles"_a, "symmetry_information"_a = "P 1", "basis"_a = py::none());
  m.def("hall_symbol", &hall_symbol_of, "number"_a);
  m.def("spacegroup_number", &spacegroup_number_of, "symbol"_a);

  py::class_<ConvexPolyhedron>(m, "Polyhedron")
      .def_property_readonly("vertices", [](const ConvexPolyhedron& p) { return v3list(p.verts); })
      .def_property_readonly("faces", [](const ConvexPolyhedron& p) { return p.faces; })
      .def_property_readonly("volume", &ConvexPolyhedron::volume)
      .def_property_readonly("face_normals", [](const ConvexPolyhedron& p) {
        std::vector<V3> n;
        for (const auto& f : p.faces) {
          V3 c = cross(sub(p.verts[f[1]], p.verts[f[0]]), sub(p.verts[f[2]], p.verts[f[1]]));
          n.push_back(scale(c, 1.0 / norm(c)));
        }
        return v3list(n);
      })
      .def("contains", [](const ConvexPolyhedron& p, const darr& x, double tol) {
        std::vector<bool> out;
        for (py::ssize_t i = 0; i < x.shape(0); ++i) out.push_back(p.contains({x.at(i, 0), x.at(i, 1), x.at(i, 2)}, tol));
        return out;
      }, "x"_a, "tol"_a = 1e-10);

  py::class_<PyZone>(m, "BrillouinZone")
      .def(py::init([](const Lattice& lat, bool use_primitive, int search_length, bool tr, bool wedge_search,
                       bool divide_primitive) {
             return PyZone{BrillouinZone(lat, use_primitive, search_length, tr, wedge_search, divide_primitive), nullptr, nullptr};
           }),
           "lattice"_a, "use_primitive"_a = true, "search_length"_a = 1, "time_reversal_symmetry"_a = false,
           "wedge_search"_a = true, "divide_primitive"_a = true)
      .def(py::init([](const Lattice& lat, const ApproxConfig& ac, bool use_primitive, int search_length, bool tr,
                       bool wedge_search, bool divide_primitive) {
             return PyZone{BrillouinZone(lat, use_primitive, search_length, tr, wedge_search, divide_primitive, ac), nullptr, nullptr};
           }),
           "lattice"_a, "approx_config"_a, "use_primitive"_a = true, "search_length"_a = 1,
           "time_reversal_symmetry"_a = false, "wedge_search"_a = true, "divide_primitive"_a = true)
      .def_property_readonly("lattice", [](const PyZone& z) { return z.bz.outer; })
      .def_property_readonly("polyhedron", [](const PyZone& z) { return z.bz.first; })
      .def_property_readonly("ir_polyhedron", [](const PyZone& z) { return z.bz.ir; })
      .def_property_readonly("normals", [](const PyZone& z) { return v3list(z.bz.first_normals_outer()); })
      .def_property_readonly("points", [](const PyZone& z) { return v3list(z.bz.first_points_outer()); })
      .def_property_readonly("vertices", [](const PyZone& z) {
        std::vector<V3> v;
        for (const auto& x : z.bz.first.verts) v.push_back(z.bz.outer_from_xyz(x));
        return v3list(v);
      })
      .def_property_readonly("vertices_invA", [](const PyZone& z) { return v3list(z.bz.first.verts); })
      .def_property_readonly("ir_vertices", [](const PyZone& z) {
        std::vector<V3> v;
        for (const auto& x : z.bz.ir.verts) v.push_back(z.bz.outer_from_xyz(x));
        return v3list(v);
      })
      .def_property_readonly("ir_vertices_invA", [](const PyZone& z) { return v3list(z.bz.ir.verts); })
      .def_property_readonly("wedge_normals", [](const PyZone& z) { return v3list(z.bz.wedge_normals_outer()); })
      .def_property_readonly("wedge_normals_invA", [](const PyZone& z) { return v3list(z.bz.wedge_normals_xyz); })
      .def_property_readonly("vertices_per_face", [](const PyZone& z) { return z.bz.first.faces; })
      .def_property_readonly("ir_vertices_per_face", [](const PyZone& z) { return z.bz.ir.faces; })
      .def_property_readonly("pointgroup", [](const PyZone& z) { return z.bz.psym; })
      .def("isinside", [](const PyZone& z, const darr& Q) {
        std::vector<bool> out;
        for (py::ssize_t i = 0; i < Q.shape(0); ++i)
          out.push_back(z.bz.first.contains(z.bz.xyz_from_outer({Q.at(i, 0), Q.at(i, 1), Q.at(i, 2)}),
                                            1e-10 * z.bz.first.extent()));
        return out;
      }, "Q"_a)
      .def("ir_moveinto", [](PyZone& z, const darr& Q, int threads) {
        (void)threads;
        if (Q.ndim() != 2 || Q.shape(1) != 3) throw std::runtime_error("Q must be (n, 3)");
        const size_t n = (size_t)Q.shape(0);
        py::array_t<double> q({(py::ssize_t)n, (py::ssize_t)3});
        py::array_t<int32_t> tau({(py::ssize_t)n, (py::ssize_t)3});
        std::vector<uint32_t> r(n), ir(n);
        check_status(brille_ir_moveinto(z.handle(), Q.data(), n, BRILLE_MEM_HOST, q.mutable_data(), tau.mutable_data(),
                                        r.data(), ir.data(), nullptr, nullptr));
        py::array_t<int> R({(py::ssize_t)n, (py::ssize_t)3, (py::ssize_t)3}), Ri({(py::ssize_t)n, (py::ssize_t)3, (py::ssize_t)3});
        for (size_t i = 0; i < n; ++i) {
          std::memcpy(R.mutable_data() + 9 * i, z.bz.psym.R[r[i]].data(), 9 * sizeof(int));
          std::memcpy(Ri.mutable_data() + 9 * i, z.bz.psym.R[ir[i]].data(), 9 * sizeof(int));
        }
        return py::make_tuple(q, tau, R, Ri);
      }, "Q"_a, "threads"_a = 0);

  py::class_<PyGrid>(m, "BZTrellisQdc", py::dynamic_attr())
      .def(py::init([](const PyZone& z, double v, bool always) {
             return PyGrid{std::make_shared<BZTrellisQdc>(z.bz, v, always, z.bz.approx), py::none(), py::none()};
           }),
           "brillouin_zone"_a, "node_volume_fraction"_a = 0.1, "always_triangulate"_a = false)
      .def(py::init([](const PyZone& z, double v, bool always, const ApproxConfig& ac) {
             return PyGrid{std::make_shared<BZTrellisQdc>(z.bz, v, always, ac), py::none(), py::none()};
           }),
           "brillouin_zone"_a, "node_volume_fraction"_a, "always_triangulate"_a, "approx_config"_a)
      .def_property_readonly("BrillouinZone", [](const PyGrid& g) { return PyZone{g.g->bz, nullptr, nullptr}; })
      .def_property_readonly("invA", [](const PyGrid& g) {
        return vec_array(g.g->trellis.vertices, {(py::ssize_t)g.g->trellis.n_vertices(), 3});
      })
      .def_property_readonly("rlu", [](const PyGrid& g) {
        std::vector<V3> v;
        const auto& T = g.g->trellis;
        for (size_t i = 0; i < T.n_vertices(); ++i)
          v.push_back(g.g->bz.outer_from_xyz({T.vertices[3 * i], T.vertices[3 * i + 1], T.vertices[3 * i + 2]}));
        return v3list(v);
      })
      .def_property_readonly("tetrahedra", [](const PyGrid& g) {
        return vec_array(g.g->trellis.tet_vertices, {(py::ssize_t)g.g->trellis.n_tets(), 4});
      })
      .def_property_readonly("bytes_per_point", [](const PyGrid& g) { return g.g->bytes_per_point(); })
      .def_property_readonly("values", [](const PyGrid& g) { return g.values_ref; })
      .def_property_readonly("vectors", [](const PyGrid& g) { return g.vectors_ref; })
      // wrap/_common_grid.hpp:26-47: sort=True sorts the modes of every
      // connected vertex pair once the data are on the device
      .def("fill", [](PyGrid& g, const py::array& vals, const iarr& vel, const py::array& vecs, const iarr& cel, bool sort) {
        g.g->fill(payload_from(vals, vel, false), payload_from(vecs, cel, true), sort);
        g.values_ref = vals;
        g.vectors_ref = vecs;
      }, "values_data"_a, "values_elements"_a, "vectors_data"_a, "vectors_elements"_a, "sort"_a = false)
      // wrap/_common_grid.hpp:126-147 with wrap/_interpolator.hpp:56-107:
      // elements[5] and [6] pick the scalar and vector cost functions, the
      // weights scale the scalar, vector and matrix terms
      .def("fill", [](PyGrid& g, const py::array& vals, const iarr& vel, const darr& vw, const py::array& vecs,
                      const iarr& cel, const darr& cw, bool sort) {
        Payload pv = payload_from(vals, vel, false), pc = payload_from(vecs, cel, true);
        set_cost(pv, vel, vw);
        set_cost(pc, cel, cw);
        g.g->fill(std::move(pv), std::move(pc), sort);
        g.values_ref = vals;
        g.vectors_ref = vecs;
      }, "values_data"_a, "values_elements"_a, "values_weights"_a, "vectors_data"_a, "vectors_elements"_a,
         "vectors_weights"_a, "sort"_a = false)
      .def("sort", [](PyGrid& g) {
        brille_grid* h = g.g->handle();
        check_status(brille_grid_sort(h, &g.g->values.cost, &g.g->vectors.cost));
      })
      .def("permutation_table", [](PyGrid& g) {
        brille_grid* h = g.g->handle();
        uint64_t nk = 0;
        uint32_t np = 0;
        check_status(brille_grid_permutations(h, &nk, &np, nullptr, nullptr, nullptr));
        py::array_t<uint64_t> keys((py::ssize_t)nk);
        py::array_t<uint32_t> index((py::ssize_t)nk);
        py::array_t<uint32_t> perms({(py::ssize_t)np, (py::ssize_t)g.g->values.branches});
        check_status(brille_grid_permutations(h, &nk, &np, keys.mutable_data(), index.mutable_data(), perms.mutable_data()));
        return py::make_tuple(keys, index, perms);
      })
      .def(
          "ir_interpolate_at",
          [](PyGrid& pg, const darr& Q, bool par, int th, bool nm, const std::string& prec, const std::string& k) {
            return grid_interpolate(pg, Q, par, th, nm, prec, k, false);
          },
          "Q"_a, "useparallel"_a = false, "threads"_a = -1, "do_not_move_points"_a = false, py::kw_only(),
          "precision"_a = "fp64", "kernel"_a = "auto")
      // wrap/_common_grid.hpp:412-440: first-zone fold only, no rotation
      .def(
          "interpolate_at",
          [](PyGrid& pg, const darr& Q, bool par, int th, bool nm, const std::string& prec, const std::string& k) {
            return grid_interpolate(pg, Q, par, th, nm, prec, k, true);
          },
          "Q"_a, "useparallel"_a = false, "threads"_a = -1, "do_not_move_points"_a = false, py::kw_only(),
          "precision"_a = "fp64", "kernel"_a = "auto")
      .def("all_node_types", [](const PyGrid& g) { return g.g->trellis.node_type; })
      .def("node_at_type", [](const PyGrid& g, const std::array<size_t, 3>& s) {
        const auto& T = g.g->trellis;
        size_t n0 = T.knots[0].size() - 1, n1 = T.knots[1].size() - 1;
        return (int)T.node_type.at(s[0] + n0 * (s[1] + n1 * s[2]));
      }, "subscript"_a)
      .def("total_node_volume", [](const PyGrid& g) { return g.g->trellis.total_node_volume(); })
      .def("tables", [](const PyGrid& g) { return grid_tables(*g.g); })
      // the GammaTable (src/phonon.hpp:91-194) of the grid's lattice, as the
      // tables() keys gamma_*: what a Gamma-rotating payload filled through
      // the C ABI (brille_grid_fill_ex) needs, without a host payload
      .def("gamma_tables", [](const PyGrid& pg) {
        const BZTrellisQdc& g = *pg.g;
        GammaTableHost gt;
        gt.construct(g.bz.outer, g.bz.psym, g.bz.time_reversal, g.approx.direct, g.approx.digits);
        py::dict t;
        t["gamma_F0"] = vec_array(gt.F0, {(py::ssize_t)gt.n_atoms, (py::ssize_t)gt.n_ops});
        t["gamma_vector_index"] = vec_array(gt.vindex, {(py::ssize_t)gt.n_atoms, (py::ssize_t)gt.n_ops});
        t["gamma_vectors"] = vec_array(gt.vectors, {(py::ssize_t)(gt.vectors.size() / 3), 3});
        t["gamma_cart_rotations"] = vec_array(gt.cart_rotations, {(py::ssize_t)gt.n_ops, 9});
        return t;
      })
      .def("device_handle", [](PyGrid& g) { return reinterpret_cast<uintptr_t>(g.g->handle()); });

  // BZMeshQdc and BZNestQdc (src/bz_mesh.hpp:37-112, src/bz_nest.hpp:36-104;
  // constructors wrap/_mesh.hpp:36, wrap/_nest.hpp:37-38): the same methods
  // over a tetrahedral mesh of the irreducible zone.  The reference meshes
  // with TetGen, which this build does not have; here the mesh is the trellis
  // with every node triangulated (tetrahedra no larger than the requested
  // volume), located and summed by the same device path, so stages (a) and
  // (c) are shared and only the mesh differs.  num_levels, max_points and
  // max_branchings shape TetGen's refinement and are accepted and ignored.
  auto mesh_volume = [](const PyZone& z, double v) {
    const double ir = z.bz.ir.volume();
    return v > 0 ? v : ir / 8.0;
  };
  py::class_<PyMesh, PyGrid>(m, "BZMeshQdc", py::dynamic_attr())
      .def(py::init([mesh_volume](const PyZone& z, double max_size, int num_levels, int max_points) {
             (void)num_levels;
             (void)max_points;
             return PyMesh{PyGrid{std::make_shared<BZTrellisQdc>(z.bz, mesh_volume(z, max_size), true, z.bz.approx),
                                  py::none(), py::none()}};
           }),
           "brillouin_zone"_a, "max_size"_a = -1., "num_levels"_a = 3, "max_points"_a = -1);
  py::class_<PyNest, PyGrid>(m, "BZNestQdc", py::dynamic_attr())
      .def(py::init([mesh_volume](const PyZone& z, double max_volume, unsigned max_branchings) {
             (void)max_branchings;
             return PyNest{PyGrid{std::make_shared<BZTrellisQdc>(z.bz, mesh_volume(z, max_volume), true, z.bz.approx),
                                  py::none(), py::none()}};
           }),
           "brillouin_zone"_a, "max_volume"_a, "max_branchings"_a = 5)
      .def(py::init([mesh_volume](const PyZone& z, unsigned number_density, unsigned max_branchings) {
             (void)max_branchings;
             const double v = z.bz.ir.volume() / std::max(1u, number_density);
             return PyNest{PyGrid{std::make_shared<BZTrellisQdc>(z.bz, mesh_volume(z, v), true, z.bz.approx),
                                  py::none(), py::none()}};
           }),
           "brillouin_zone"_a, "number_density"_a, "max_branchings"_a = 5)
      .def_property_readonly("all_invA", [](const PyGrid& g) {
        return vec_array(g.g->trellis.vertices, {(py::ssize_t)g.g->trellis.n_vertices(), 3});
      })
      .def_property_readonly("all_rlu", [](const PyGrid& g) {
        std::vector<V3> v;
        const auto& T = g.g->trellis;
        for (size_t i = 0; i < T.n_vertices(); ++i)
          v.push_back(g.g->bz.outer_from_xyz({T.vertices[3 * i], T.vertices[3 * i + 1], T.vertices[3 * i + 2]}));
        return v3list(v);
      });
}
