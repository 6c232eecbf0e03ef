#include "grid.hpp"

#include <cfloat>
#include <cstring>
#include <stdexcept>

namespace brille_amd {

std::string last_error_text() {
  char buf[1024];
  brille_last_error(buf, sizeof(buf));
  return buf;
}

void check_status(int rc) {
  if (rc != BRILLE_OK) throw std::runtime_error(last_error_text());
}

DeviceGrid::~DeviceGrid() {
  if (h) brille_grid_destroy(h);
}

// ---------------------------------------------------------------------------
void GammaTableHost::construct(const Lattice& lat, const PointSymmetry& ps, bool time_reversal, double e_tol, int n_tol) {
  Symmetry spg = lat.space;
  if (time_reversal) spg = spg.add_space_inversion();
  const Basis& bs = lat.basis;
  n_atoms = static_cast<uint32_t>(bs.size());
  n_ops = static_cast<uint32_t>(ps.size());
  F0.assign(static_cast<size_t>(n_atoms) * n_ops, 0);
  vindex.assign(static_cast<size_t>(n_atoms) * n_ops, 0);
  vectors.assign(3, 0.0);  // (0,0,0) first
  std::vector<size_t> p2s(n_ops);
  for (size_t i = 0; i < n_ops; ++i) {
    p2s[i] = spg.find_matrix_index(ps.R[i]);
    if (p2s[i] >= spg.size())
      throw std::runtime_error("Something has gone wrong with the correspondence of spacegroup to pointgroup");
  }
  double Trel = DBL_EPSILON * (static_cast<double>(n_tol) * 10000.0), Tabs = 5e-15;
  if (e_tol > Trel) Trel = e_tol;
  if (e_tol > Tabs) Tabs = e_tol;
  for (uint32_t k = 0; k < n_atoms; ++k)
    for (uint32_t r = 0; r < n_ops; ++r) {
      const Motion& motion = spg.ops[p2s[r]];
      uint32_t l = 0;
      if (!bs.equivalent_after_operation(k, motion, e_tol, n_tol, &l))
        throw std::runtime_error("All atoms in the basis *must* be mapped to an equivalent atom for *all* symmetry operations");
      V3 vec = motion.inverse().move_point(bs.positions[l]);
      for (int i = 0; i < 3; ++i) vec[i] -= bs.positions[k][i];
      uint32_t count = static_cast<uint32_t>(vectors.size() / 3), v = count;
      for (uint32_t c = 0; c < count; ++c) {
        double dx = vectors[3 * c] - vec[0], dy = vectors[3 * c + 1] - vec[1], dz = vectors[3 * c + 2] - vec[2];
        double nrm = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (nrm <= Tabs + Trel * nrm || nrm < DBL_MIN) { v = c; break; }
      }
      if (v == count) vectors.insert(vectors.end(), vec.begin(), vec.end());
      F0[static_cast<size_t>(k) * n_ops + r] = l;
      vindex[static_cast<size_t>(k) * n_ops + r] = v;
    }
  // Cartesian rotations A R A^-1  src/interpolator.hpp:411-420
  cart_rotations.resize(static_cast<size_t>(n_ops) * 9);
  M3 from_xyz = lat.recip;
  for (auto& x : from_xyz) x /= kTwoPi;
  from_xyz = transpose(from_xyz);
  for (uint32_t j = 0; j < n_ops; ++j) {
    M3 t = mat_mul(lat.real, to_double(ps.R[j]));
    M3 c = mat_mul(t, from_xyz);
    std::memcpy(&cart_rotations[9 * static_cast<size_t>(j)], c.data(), 9 * sizeof(double));
  }
}

// ---------------------------------------------------------------------------
void ZoneTables::build(const BrillouinZone& bz) {
  bz.first_planes_outer(outer_a, outer_b, outer_c);
  const size_t F = bz.first.faces.size();
  prim_a.resize(3 * F);
  prim_b.resize(3 * F);
  prim_c.resize(3 * F);
  for (size_t f = 0; f < F; ++f) {
    V3 a{outer_a[3 * f], outer_a[3 * f + 1], outer_a[3 * f + 2]};
    V3 b{outer_b[3 * f], outer_b[3 * f + 1], outer_b[3 * f + 2]};
    V3 c{outer_c[3 * f], outer_c[3 * f + 1], outer_c[3 * f + 2]};
    V3 pa = bz.prim_from_outer(a), pb = bz.prim_from_outer(b), pc = bz.prim_from_outer(c);
    for (int i = 0; i < 3; ++i) { prim_a[3 * f + i] = pa[i]; prim_b[3 * f + i] = pb[i]; prim_c[3 * f + i] = pc[i]; }
  }
  // moveinto's normals, taus and tau lengths  src/bz_move.cpp:137-140
  const M3& Gp = bz.inner.recip_metric;
  auto mnorm = [&](const V3& x) {
    V3 t = mat_vec(Gp, x);
    return std::sqrt(t[0] * x[0] + t[1] * x[1] + t[2] * x[2]);
  };
  prim_n.clear();
  for (const auto& n : bz.first_normals_outer()) {
    V3 pn = bz.prim_from_outer(n);
    double l = mnorm(pn);
    for (int i = 0; i < 3; ++i) prim_n.push_back(pn[i] / l);
  }
  prim_taus.clear();
  prim_tau_lens.clear();
  for (const auto& t : bz.first_taus) {
    for (int i = 0; i < 3; ++i) prim_taus.push_back(t[i]);
    prim_tau_lens.push_back(mnorm(V3{double(t[0]), double(t[1]), double(t[2])}));
  }
  wedge.clear();
  for (const auto& n : bz.wedge_normals_outer()) wedge.insert(wedge.end(), n.begin(), n.end());
  rotations.clear();
  for (const auto& R : bz.psym.R) rotations.insert(rotations.end(), R.begin(), R.end());
  inverse = bz.psym.inverse_table();
}

void ZoneTables::describe(const BrillouinZone& bz, brille_grid_desc& d) const {
  std::memcpy(d.outer_to_xyz, bz.outer.recip.data(), sizeof(d.outer_to_xyz));
  std::memcpy(d.outer_metric_recip, bz.outer.recip_metric.data(), sizeof(d.outer_metric_recip));
  std::memcpy(d.outer_metric_real, bz.outer.real_metric.data(), sizeof(d.outer_metric_real));
  d.outer_cross_scale = bz.outer.volume_recip() / kTwoPi;
  d.transform_needed = bz.transform_needed ? 1 : 0;
  IM3 P6t = transpose(primitive_6P(bz.outer.bravais));
  IM3 invPt = transpose(primitive_invP(bz.outer.bravais));
  std::memcpy(d.to_primitive_6Pt, P6t.data(), sizeof(d.to_primitive_6Pt));
  std::memcpy(d.from_primitive_invPt, invPt.data(), sizeof(d.from_primitive_invPt));
  std::memcpy(d.prim_metric_recip, bz.inner.recip_metric.data(), sizeof(d.prim_metric_recip));
  std::memcpy(d.prim_metric_real, bz.inner.real_metric.data(), sizeof(d.prim_metric_real));
  d.prim_cross_scale = bz.inner.volume_recip() / kTwoPi;
  d.n_faces = static_cast<uint32_t>(bz.first.faces.size());
  d.prim_plane_a = prim_a.data();
  d.prim_plane_b = prim_b.data();
  d.prim_plane_c = prim_c.data();
  d.prim_normals = prim_n.data();
  d.prim_taus = prim_taus.data();
  d.prim_tau_lens = prim_tau_lens.data();
  d.outer_plane_a = outer_a.data();
  d.outer_plane_b = outer_b.data();
  d.outer_plane_c = outer_c.data();
  d.n_wedge = static_cast<uint32_t>(wedge.size() / 3);
  d.wedge_normals = wedge.empty() ? nullptr : wedge.data();
  d.no_ir_mirroring = 1;
  d.has_wedge_seed = bz.has_seed ? 1 : 0;
  V3 s = bz.outer_from_xyz(bz.seed_xyz);
  for (int i = 0; i < 3; ++i) d.wedge_seed[i] = s[i];
  d.n_ops = static_cast<uint32_t>(bz.psym.size());
  d.rotations = rotations.data();
  d.inverse_index = inverse.data();
  d.identity_index = static_cast<uint32_t>(bz.psym.find_identity_index());
  d.float_tolerance = bz.approx.reciprocal;
  d.approx_digits = bz.approx.digits;
}

std::shared_ptr<DeviceGrid> zone_handle(const BrillouinZone& bz, ZoneTables& store) {
  store.build(bz);
  brille_grid_desc d;
  std::memset(&d, 0, sizeof(d));
  store.describe(bz, d);
  auto dev = std::make_shared<DeviceGrid>();
  check_status(brille_grid_create(&d, &dev->h));
  return dev;
}

// ---------------------------------------------------------------------------
void parse_payload(Payload& p, const std::vector<size_t>& shape, const std::vector<int>& el, bool complex_data) {
  if (shape.empty()) throw std::runtime_error("Interpolator data is expected to be 1- to 5-D");
  p.shape = shape;
  for (int i = 0; i < 3; ++i) p.elements[i] = 0;
  for (size_t i = 0; i < el.size() && i < 3; ++i) p.elements[i] = static_cast<uint32_t>(el[i]);
  p.rot = BRILLE_ROT_VECTOR;
  if (el.size() > 3) {
    switch (el[3]) {
      case 2: p.rot = BRILLE_ROT_GAMMA; break;
      case 1: p.rot = BRILLE_ROT_PSEUDOVECTOR; break;
      case 0: p.rot = BRILLE_ROT_VECTOR; break;
      default: throw std::runtime_error("Unknown RotatesLike value " + std::to_string(el[3]));
    }
  }
  p.lu = BRILLE_LU_REAL_LATTICE;
  if (el.size() > 4) {
    switch (el[4]) {
      case 4: p.lu = BRILLE_LU_RECIPROCAL_LATTICE; break;
      case 3: p.lu = BRILLE_LU_REAL_LATTICE; break;
      case 2: p.lu = BRILLE_LU_INVERSE_ANGSTROM; break;
      case 1: p.lu = BRILLE_LU_ANGSTROM; break;
      case 0: p.lu = BRILLE_LU_NONE; break;
      default: throw std::runtime_error("Unknown LengthUnit value " + std::to_string(el[4]));
    }
  }
  if (p.elements[1] % 3) throw std::logic_error("Vectors must have 3N elements per branch");
  if (p.elements[2] % 9) throw std::logic_error("Matrices must have 9N elements per branch");
  uint32_t x = p.span();
  const size_t nd = shape.size();
  switch (nd) {
    case 1:
      if (x == 0) x = p.elements[0] = 1;
      if (x > 1) throw std::runtime_error("1-D data must represent one scalar per point!");
      break;
    case 2:
      if (x == 0) x = p.elements[0] = static_cast<uint32_t>(shape[1]);
      if (shape[1] % x) throw std::runtime_error("2-D data requires an integer number of branches!");
      break;
    case 3:
      if (x == 0) x = p.elements[0] = static_cast<uint32_t>(shape[2]);
      if (shape[2] != x) throw std::runtime_error("3-D data requires that the last dimension matches the specified number of elements!");
      break;
    case 4:
      if (shape[3] != 3) throw std::runtime_error("4-D data can only be 3-vectors");
      if (x == 0) x = p.elements[1] = static_cast<uint32_t>(shape[2] * 3);
      if (shape[2] * 3 != x) throw std::runtime_error("4-D data requires that the last two dimensions match the specified number of vector elements!");
      break;
    case 5:
      if (shape[3] != 3 || shape[4] != 3) throw std::runtime_error("5-D data can only be matrices");
      if (x == 0) x = p.elements[2] = static_cast<uint32_t>(shape[2] * 9);
      if (shape[2] * 9 != x) throw std::runtime_error("5-D data requires the last three dimensions match the specified number of matrix elements!");
      break;
    default:
      throw std::runtime_error("Interpolator data is expected to be 1- to 5-D");
  }
  // branches  src/interpolator.hpp:309-332
  uint32_t b = nd > 1 ? static_cast<uint32_t>(shape[1]) : 1u;
  if (nd == 2) {
    uint32_t y = p.span();
    if (y > 0) b /= y;
  }
  p.branches = b;
  (void)complex_data;
}

// ---------------------------------------------------------------------------
BZTrellisQdc::BZTrellisQdc(const BrillouinZone& b, double node_volume, bool always_triangulate, ApproxConfig cfg)
    : bz(b), approx(cfg) {
  // node_volume_fraction is forwarded as the absolute maximum node volume
  // (wrap/_trellis.hpp:37 -> src/bz_trellis.hpp:56-59 -> src/trellis_poly.hpp:154-161)
  trellis = build_trellis(bz.ir, node_volume, always_triangulate);
  zone.build(bz);
}

void BZTrellisQdc::fill(Payload v, Payload c, bool sort) {
  std::lock_guard<std::mutex> lock(*host_mu);
  const size_t V = trellis.n_vertices();
  if (v.shape[0] != V || c.shape[0] != V)
    throw std::runtime_error("Provided " + std::to_string(v.shape[0]) + " arrays but " + std::to_string(V) + " were expected!");
  if (v.branches != c.branches)
    throw std::runtime_error("The values and vectors must have matching number of points and branches");
  values = std::move(v);
  vectors = std::move(c);
  filled = true;
  device_filled = false;
  sort_requested = sort;
  if (vectors.rot == BRILLE_ROT_GAMMA)
    gamma.construct(bz.outer, bz.psym, bz.time_reversal, approx.direct, approx.digits);
}

brille_grid_desc BZTrellisQdc::describe() const {
  brille_grid_desc d;
  std::memset(&d, 0, sizeof(d));
  zone.describe(bz, d);
  for (int i = 0; i < 3; ++i) {
    d.n_knots[i] = static_cast<uint32_t>(trellis.knots[i].size());
    d.knots[i] = trellis.knots[i].data();
  }
  d.n_nodes = trellis.n_nodes();
  d.node_type = trellis.node_type.data();
  d.node_index = trellis.node_index.data();
  d.n_cubes = trellis.n_cubes();
  d.cube_vertices = trellis.cube_vertices.data();
  d.n_polys = trellis.n_polys();
  d.poly_tet_offset = trellis.poly_offset.data();
  d.n_tets = trellis.n_tets();
  d.tet_vertices = trellis.tet_vertices.data();
  d.tet_circumsphere = trellis.tet_circum.data();
  d.tet_volume = trellis.tet_volume.data();
  d.n_vertices = trellis.n_vertices();
  d.vertices = trellis.vertices.data();
  return d;
}

brille_grid* BZTrellisQdc::handle() {
  std::lock_guard<std::mutex> lock(*host_mu);
  if (!device) {
    brille_grid_desc d = describe();
    device = std::make_shared<DeviceGrid>();
    check_status(brille_grid_create(&d, &device->h));
    device_filled = false;
  }
  if (!filled) throw std::runtime_error("The interpolation data must be filled before interpolating.");
  if (!device_filled) {
    brille_data_desc dv{}, dc{};
    dv.n_vertices = values.shape[0];
    dv.branches = values.branches;
    for (int i = 0; i < 3; ++i) { dv.elements[i] = values.elements[i]; dc.elements[i] = vectors.elements[i]; }
    dv.rotates_like = values.rot;
    dv.length_unit = values.lu;
    dv.is_complex = 0;
    dv.data = values.data.data();
    dc.n_vertices = vectors.shape[0];
    dc.branches = vectors.branches;
    dc.rotates_like = vectors.rot;
    dc.length_unit = vectors.lu;
    dc.is_complex = 1;
    dc.data = vectors.data.data();
    brille_gamma_desc g{};
    const brille_gamma_desc* gp = nullptr;
    if (vectors.rot == BRILLE_ROT_GAMMA) {
      g.n_atoms = gamma.n_atoms;
      g.n_ops = gamma.n_ops;
      g.F0 = gamma.F0.data();
      g.vector_index = gamma.vindex.data();
      g.n_vectors = static_cast<uint32_t>(gamma.vectors.size() / 3);
      g.vectors = gamma.vectors.data();
      g.cart_rotations = gamma.cart_rotations.data();
      gp = &g;
    }
    check_status(brille_grid_fill(device->h, &dv, &dc, gp, nullptr));
    // wrap/_common_grid.hpp:39-43: sort after the data are in place
    if (sort_requested) check_status(brille_grid_sort(device->h, &values.cost, &vectors.cost));
    device_filled = true;
  }
  return device->h;
}

uint64_t BZTrellisQdc::bytes_per_point(int precision) const {
  const uint64_t s = precision == BRILLE_FP32 ? 4 : 8;
  return static_cast<uint64_t>(values.branches) * values.span() * s +
         static_cast<uint64_t>(vectors.branches) * vectors.span() * 2 * s;
}

void BZTrellisQdc::ir_interpolate_at(const double* Q, uint64_t n, bool no_move, int precision, int kernel, void* vals,
                                     void* vecs, const brille_aux_out* aux, bool first_zone) {
  brille_grid* h = handle();
  brille_run_opts o{};
  o.no_move = no_move ? 1 : 0;
  o.precision = precision;
  o.memory = BRILLE_MEM_HOST;
  o.kernel = kernel;
  o.check = 1;
  if (first_zone) check_status(brille_interpolate_at(h, Q, n, &o, vals, vecs, aux, nullptr));
  else check_status(brille_ir_interpolate_at(h, Q, n, &o, vals, vecs, aux, nullptr));
}

}  // namespace brille_amd
