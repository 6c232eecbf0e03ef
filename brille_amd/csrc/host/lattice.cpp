#include "lattice.hpp"

#include <stdexcept>

namespace brille_amd {

// The conventional-to-primitive transformations of International Tables
// Vol. A (centring-dependent P, stored as 6P to stay integer), in the same
// matrix convention the reference uses (src/primitive.hpp:112-155).
IM3 primitive_6P(Bravais b) {
  switch (b) {
    case Bravais::I: return {-3, 3, 3, 3, -3, 3, 3, 3, -3};
    case Bravais::F: return {0, 3, 3, 3, 0, 3, 3, 3, 0};
    case Bravais::A: return {6, 0, 0, 0, 3, -3, 0, 3, 3};
    case Bravais::B: return {3, 0, -3, 0, 6, 0, 3, 0, 3};
    case Bravais::C: return {3, 3, 0, -3, 3, 0, 0, 0, 6};
    case Bravais::R: return {4, -2, -2, 2, 2, -4, 2, 2, 2};
    case Bravais::_: throw std::runtime_error("Invalid Bravais centring");
    default: return {6, 0, 0, 0, 6, 0, 0, 0, 6};
  }
}

IM3 primitive_invP(Bravais b) {
  switch (b) {
    case Bravais::I: return {0, 1, 1, 1, 0, 1, 1, 1, 0};
    case Bravais::F: return {-1, 1, 1, 1, -1, 1, 1, 1, -1};
    case Bravais::A: return {1, 0, 0, 0, 1, 1, 0, -1, 1};
    case Bravais::B: return {1, 0, 1, 0, 1, 0, -1, 0, 1};
    case Bravais::C: return {1, -1, 0, 1, 1, 0, 0, 0, 1};
    case Bravais::R: return {1, 0, 1, -1, 1, 1, 0, -1, 1};
    case Bravais::_: throw std::runtime_error("Invalid Bravais centring");
    default: return {1, 0, 0, 0, 1, 0, 0, 0, 1};
  }
}

bool primitive_does_anything(Bravais b) { return b != Bravais::P && b != Bravais::_; }

namespace {
// exact cosine/sine of the angles crystallographers write down
void cos_sin_deg(double d, double* c, double* s) {
  const double h3 = 0.86602540378443864676372317075294, h2 = 0.70710678118654752440084436210485;
  if (d == 0) { *c = 1; *s = 0; return; }
  if (d == 30) { *c = h3; *s = 0.5; return; }
  if (d == 45) { *c = h2; *s = h2; return; }
  if (d == 60) { *c = 0.5; *s = h3; return; }
  if (d == 90) { *c = 0; *s = 1; return; }
  if (d == 120) { *c = -0.5; *s = h3; return; }
  if (d == 135) { *c = -h2; *s = h2; return; }
  if (d == 150) { *c = -h3; *s = 0.5; return; }
  if (d == 180) { *c = -1; *s = 0; return; }
  double r = d * kPi / 180.0;
  *c = std::cos(r);
  *s = std::sin(r);
}

M3 metric_of(const M3& columns) { return mat_mul(transpose(columns), columns); }
}  // namespace

void Lattice::set_from_recip(const M3& B) {
  recip = B;
  M3 inv = inverse(B);
  for (auto& x : inv) x *= kTwoPi;
  real = transpose(inv);
  real_metric = metric_of(real);
  recip_metric = metric_of(recip);
}

void Lattice::set_from_real(const M3& A) {
  real = A;
  M3 inv = inverse(A);
  for (auto& x : inv) x *= kTwoPi;
  recip = transpose(inv);
  real_metric = metric_of(real);
  recip_metric = metric_of(recip);
}

Lattice Lattice::from_parameters(const V3& lengths, const V3& angles, const Symmetry& s, Bravais b,
                                 const Basis& basis, bool real_space) {
  Lattice L;
  L.space = s;
  L.bravais = b;
  L.point = PointSymmetry::from_spacegroup(s);
  L.basis = basis;
  bool degrees = std::max(angles[0], std::max(angles[1], angles[2])) >= kPi;
  V3 c{}, sn{};
  for (int i = 0; i < 3; ++i) {
    if (degrees) cos_sin_deg(angles[i], &c[i], &sn[i]);
    else { c[i] = std::cos(angles[i]); sn[i] = std::sin(angles[i]); }
  }
  // dual_lattice_parameters  src/lattice_dual.hpp:73-88
  double cos_sum = 0, cos_prod = 2;
  for (int i = 0; i < 3; ++i) { cos_sum += c[i] * c[i]; cos_prod *= c[i]; }
  double uv2 = 1.0 - cos_sum + cos_prod;
  if (uv2 <= 0) throw std::invalid_argument("lattice angles give no volume");
  double uv = std::sqrt(uv2);
  V3 dl{}, dc{}, ds{};
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    dl[i] = kTwoPi * sn[i] / lengths[i] / uv;
    dc[i] = (c[j] * c[k] - c[i]) / (sn[j] * sn[k]);
    ds[i] = std::sqrt(1.0 - dc[i] * dc[i]);
  }
  V3 rv, rc, rs, dv, dcos;
  if (real_space) { dv = lengths; dcos = c; rv = dl; rc = dc; rs = ds; }
  else { rv = lengths; rc = c; rs = sn; dv = dl; dcos = dc; }
  // Busing & Levy B matrix with the 2 pi convention  src/lattice_dual.hpp:493-500
  M3 B{rv[0], rv[1] * rc[2], rv[2] * rc[1],
       0.0, rv[1] * rs[2], -rv[2] * rs[1] * dcos[0],
       0.0, 0.0, kTwoPi / dv[2]};
  L.set_from_recip(B);
  return L;
}

Lattice Lattice::from_vectors(const M3& v, bool row_vectors, const Symmetry& s, Bravais b,
                              const Basis& basis, bool real_space) {
  Lattice L;
  L.space = s;
  L.bravais = b;
  L.point = PointSymmetry::from_spacegroup(s);
  L.basis = basis;
  M3 cols = row_vectors ? transpose(v) : v;
  if (real_space) L.set_from_real(cols);
  else L.set_from_recip(cols);
  return L;
}

V3 Lattice::lengths(bool real_space) const {
  const M3& m = real_space ? real : recip;
  V3 out{};
  for (int j = 0; j < 3; ++j) out[j] = std::sqrt(m[j] * m[j] + m[3 + j] * m[3 + j] + m[6 + j] * m[6 + j]);
  return out;
}

V3 Lattice::angles(bool real_space) const {
  const M3& m = real_space ? real : recip;
  V3 out{};
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    V3 a{m[j], m[3 + j], m[6 + j]}, b{m[k], m[3 + k], m[6 + k]};
    out[i] = std::acos(dot(a, b) / norm(a) / norm(b));
  }
  return out;
}

Lattice Lattice::primitive() const {
  if (!primitive_does_anything(bravais)) return *this;
  Lattice P(*this);
  // Impl::primitive  src/lattice_dual.hpp:736-771
  M3 pv = mat_mul(real, to_double(primitive_6P(bravais)));
  for (auto& x : pv) x /= 6.0;
  M3 pr = mat_mul(recip, to_double(transpose(primitive_invP(bravais))));
  P.real = pv;
  P.recip = pr;
  P.real_metric = metric_of(pv);
  P.recip_metric = metric_of(pr);
  P.bravais = Bravais::P;
  return P;
}

bool Lattice::is_triclinic() const {
  for (const auto& r : point.R)
    if (rotation_order(r) > 1 && isometry_value(r) != -1) return false;
  return true;
}

}  // namespace brille_amd
