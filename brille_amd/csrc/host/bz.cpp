#include "bz.hpp"

#include <algorithm>
#include <stdexcept>

namespace brille_amd {

BrillouinZone::BrillouinZone(const Lattice& lat, bool use_primitive, int search_length, bool tr,
                             bool wedge_search, bool /*divide_primitive*/, ApproxConfig cfg)
    : outer(lat), inner(use_primitive ? lat.primitive() : lat), time_reversal(tr), approx(cfg) {
  psym = outer.point;
  if (time_reversal) psym = psym.add_space_inversion();
  has_inversion = time_reversal || outer.has_space_inversion();
  transform_needed = use_primitive && primitive_does_anything(outer.bravais);
  voro_search(std::max(search_length, 1) + 2);
  double expect = std::abs(inner.volume_recip());
  double got = first.volume();
  if (!(std::abs(got - expect) <= 1e-9 * expect))
    throw std::runtime_error("failed to produce a first Brillouin zone with the primitive cell volume");
  ir = first;
  if (wedge_search && psym.size() > 1) build_wedge();
  double goal = first.volume() / static_cast<double>(wedge_search ? psym.size() : 1);
  if (wedge_search && !(std::abs(ir.volume() - goal) <= 1e-8 * goal))
    throw std::runtime_error("Failed to find an irreducible Brillouin zone.");
}

V3 BrillouinZone::prim_from_outer(const V3& hkl) const {
  if (!transform_needed) return hkl;
  IM3 P6t = transpose(primitive_6P(outer.bravais));
  V3 out{};
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += static_cast<double>(P6t[3 * i + k]) * hkl[k];
    out[i] = s / 6.0;
  }
  return out;
}

void BrillouinZone::voro_search(int extent) {
  // candidate lattice vectors of the primitive reciprocal lattice
  struct Cand { I3 ijk; V3 x; double len; };
  std::vector<Cand> cand;
  for (int i = -extent; i <= extent; ++i)
    for (int j = -extent; j <= extent; ++j)
      for (int k = -extent; k <= extent; ++k) {
        if (!i && !j && !k) continue;
        V3 h{double(i), double(j), double(k)};
        V3 x = mat_vec(inner.recip, h);
        cand.push_back({I3{i, j, k}, x, norm(x)});
      }
  std::stable_sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.len < b.len; });
  // starting box bounding every nearest-neighbour bisector
  double R = 0;
  for (const auto& c : cand) R = std::max(R, c.len);
  R = 0.0;
  for (int j = 0; j < 3; ++j) R += norm(V3{inner.recip[j], inner.recip[3 + j], inner.recip[6 + j]});
  ConvexPolyhedron P = ConvexPolyhedron::box(V3{-R, -R, -R}, V3{R, R, R});
  const double tol = 1e-12 * R;
  std::vector<I3> tags;
  for (size_t ci = 0; ci < cand.size(); ++ci) {
    const auto& c = cand[ci];
    double vmax = 0;
    for (const auto& v : P.verts) vmax = std::max(vmax, norm(v));
    if (c.len / 2.0 > vmax + tol) break;
    Plane pl;
    pl.n = c.x;
    pl.d = c.len * c.len / 2.0;
    pl.tag = static_cast<int>(ci);
    P = P.clip(pl, tol);
    if (P.empty()) throw std::runtime_error("first Brillouin zone search collapsed");
  }
  for (const auto& pl : P.planes)
    if (pl.tag < 0) throw std::runtime_error("first Brillouin zone search did not converge; increase search_length");
  first = P;
  first_taus.clear();
  for (const auto& pl : P.planes) first_taus.push_back(cand[static_cast<size_t>(pl.tag)].ijk);
}

void BrillouinZone::build_wedge() {
  // reciprocal-space action of each operation on conventional rlu: q -> R^T q
  const M3 B = outer.recip, Binv = inverse(outer.recip);
  const double golden = 0.61803398874989484820;
  const std::vector<V3> seeds{{1.0, golden, golden * golden},
                              {1.0, 0.5 * golden, 0.25 * golden},
                              {0.7548776662466927, 1.0, 0.5698402909980532},
                              {0.3, 0.9, 0.2}};
  const size_t e = psym.find_identity_index();
  for (const auto& s : seeds) {
    V3 p0x = s;
    V3 p0 = mat_vec(Binv, p0x);
    std::vector<V3> normals;  // xyz, unnormalised p0 - R^T p0
    bool generic = true;
    for (size_t j = 0; j < psym.size(); ++j) {
      if (j == e) continue;
      V3 pj = mat_vec(to_double(transpose(psym.R[j])), p0);
      V3 d = sub(mat_vec(B, p0), mat_vec(B, pj));
      if (norm(d) < 1e-6 * norm(p0x)) { generic = false; break; }
      normals.push_back(d);
    }
    if (!generic) continue;
    ConvexPolyhedron P = first;
    const double tol = 1e-12 * first.extent();
    for (size_t k = 0; k < normals.size(); ++k) {
      Plane pl;
      V3 n = scale(normals[k], 1.0 / norm(normals[k]));
      pl.n = scale(n, -1.0);  // inside: n.x >= 0
      pl.d = 0.0;
      pl.tag = 100000 + static_cast<int>(k);
      P = P.clip(pl, tol);
      if (P.empty()) break;
    }
    if (P.empty()) continue;
    ir = P;
    wedge_normals_xyz.clear();
    std::vector<int> used;
    for (const auto& pl : P.planes)
      if (pl.tag >= 100000 && std::find(used.begin(), used.end(), pl.tag) == used.end()) used.push_back(pl.tag);
    std::sort(used.begin(), used.end());
    for (int t : used) {
      V3 n = normals[static_cast<size_t>(t - 100000)];
      wedge_normals_xyz.push_back(scale(n, 1.0 / norm(n)));
    }
    seed_xyz = p0x;
    has_seed = true;
    wedge_searched = true;
    return;
  }
  throw std::runtime_error("Failed to find an irreducible Brillouin zone: no generic seed point");
}

void BrillouinZone::first_planes_outer(std::vector<double>& a, std::vector<double>& b, std::vector<double>& c) const {
  a.clear(); b.clear(); c.clear();
  for (const auto& f : first.faces) {
    V3 pa = outer_from_xyz(first.verts[f[0]]), pb = outer_from_xyz(first.verts[f[1]]), pc = outer_from_xyz(first.verts[f[2]]);
    a.insert(a.end(), pa.begin(), pa.end());
    b.insert(b.end(), pb.begin(), pb.end());
    c.insert(c.end(), pc.begin(), pc.end());
  }
}

std::vector<V3> BrillouinZone::first_normals_outer() const {
  std::vector<V3> out;
  for (const auto& f : first.faces) {
    const V3& A = first.verts[f[0]];
    const V3& Bv = first.verts[f[1]];
    const V3& C = first.verts[f[2]];
    V3 n = cross(sub(Bv, A), sub(C, Bv));  // three_point_normal  src/geometry.hpp:249-252
    n = scale(n, 1.0 / norm(n));
    out.push_back(outer_from_xyz(n));
  }
  return out;
}

std::vector<V3> BrillouinZone::first_points_outer() const {
  std::vector<V3> out;
  for (const auto& f : first.faces) {
    V3 s{0, 0, 0};
    for (auto i : f) s = add(s, first.verts[i]);
    out.push_back(outer_from_xyz(scale(s, 1.0 / static_cast<double>(f.size()))));
  }
  return out;
}

std::vector<V3> BrillouinZone::wedge_normals_outer() const {
  std::vector<V3> out;
  for (const auto& n : wedge_normals_xyz) out.push_back(outer_from_xyz(n));
  return out;
}

}  // namespace brille_amd
