// Host-side precomputation of the fast fold-in tables (see fast_fold.hpp).
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/brille_amd.h"

namespace brille_amd {

struct FastTables {
  std::vector<double> prim_aff, outer_aff, orbit;
  double aff_margin = 0, orbit_margin = 0;
  bool has_orbit = false;
};

// pseudo_orient3d(a,b,c,x) in a lattice with reciprocal volume V equals
// V det[a-x, b-x, c-x] on the components = V det[a,b,c] - x . V (a x b + b x c + c x a).
inline void face_affine_forms(const double* A, const double* B, const double* C, uint32_t F, double vol,
                              std::vector<double>& out, double* scale) {
  out.assign(4 * (size_t)F, 0.0);
  double S = 0;
  for (uint32_t j = 0; j < F; ++j) {
    const double* a = A + 3 * j;
    const double* b = B + 3 * j;
    const double* c = C + 3 * j;
    double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    double l1 = 0;
    for (int k = 0; k < 3; ++k) {
      out[4 * j + k] = vol * (ab[k] + bc[k] + ca[k]);
      l1 += std::abs(out[4 * j + k]);
    }
    out[4 * j + 3] = vol * det;
    S = std::max(S, std::abs(out[4 * j + 3]) + 4.0 * l1);
  }
  *scale = S;
}

// approx_float tolerances (src/approx_float.hpp:79-100) and the half-width of
// the band around zero that the tolerant compare accepts: |x| <= Tabs + Trel |x|
inline double zero_band(double Ttol, int digits, double* Trel_out) {
  double Trel = 2.220446049250313e-16 * (double)digits * 10000.0, Tabs = 5e-15;
  if (Ttol > Trel) Trel = Ttol;
  if (Ttol > Tabs) Tabs = Ttol;
  *Trel_out = Trel;
  return Trel < 1.0 ? Tabs / (1.0 - Trel) : 1e300;
}

// Largest score lead a point-group operation other than the top scorer can
// trail by and still pass a wedge test that accepts every facet value down to
// -1.  The score of operation j is q . u_j (u_j = R_j G* p0) and its wedge test
// reads q . c_jk >= -T with c_jk = R_j G* n_k for the wedge normals n_k.
// Writing u_j - u_i = sum_k l_k c_jk with l_k >= 0 gives
// q . (u_i - u_j) <= T sum_k l_k whenever operation j passes: the lead of the
// top scorer i over any j that passes is at most T times that sum.  Each
// difference is written with at most three facets (a 3D cone needs no more);
// the smallest sum found is kept per pair and the largest over all pairs
// returned, or -1 when some difference has no such form (the fast test is
// then not used).
inline double wedge_lead_bound(const brille_grid_desc* d, const std::vector<double>& u) {
  const uint32_t G = d->n_ops, W = d->n_wedge;
  const double* M = d->outer_metric_recip;
  std::vector<double> c(3 * (size_t)W * G);
  for (uint32_t j = 0; j < G; ++j)
    for (uint32_t k = 0; k < W; ++k) {
      const double* n = d->wedge_normals + 3 * k;
      double Gn[3];
      for (int a = 0; a < 3; ++a) Gn[a] = M[3 * a] * n[0] + M[3 * a + 1] * n[1] + M[3 * a + 2] * n[2];
      const int32_t* R = d->rotations + 9 * j;
      for (int a = 0; a < 3; ++a)
        c[3 * ((size_t)j * W + k) + a] = R[3 * a] * Gn[0] + R[3 * a + 1] * Gn[1] + R[3 * a + 2] * Gn[2];
    }
  auto solve3 = [](const double* a, const double* b, const double* e, const double* v, double* l) {
    // l with l0 a + l1 b + l2 e = v (Cramer); false when singular
    const double det = a[0] * (b[1] * e[2] - b[2] * e[1]) - b[0] * (a[1] * e[2] - a[2] * e[1]) +
                       e[0] * (a[1] * b[2] - a[2] * b[1]);
    const double na = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double nb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const double ne = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    if (std::abs(det) <= 1e-9 * na * nb * ne) return false;
    auto dt = [](const double* x, const double* y, const double* z) {
      return x[0] * (y[1] * z[2] - y[2] * z[1]) - y[0] * (x[1] * z[2] - x[2] * z[1]) + z[0] * (x[1] * y[2] - x[2] * y[1]);
    };
    l[0] = dt(v, b, e) / det;
    l[1] = dt(a, v, e) / det;
    l[2] = dt(a, b, v) / det;
    return true;
  };
  double worst = 0;
  for (uint32_t j = 0; j < G; ++j) {
    const double* cj = c.data() + 3 * (size_t)j * W;
    for (uint32_t i = 0; i < G; ++i) {
      if (i == j) continue;
      double v[3];
      for (int a = 0; a < 3; ++a) v[a] = u[3 * j + a] - u[3 * i + a];
      const double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (nv == 0) continue;
      double best = 1e300;
      // one facet, two facets (in their plane), three facets
      for (uint32_t k = 0; k < W; ++k) {
        const double* a = cj + 3 * k;
        const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const double l = (v[0] * a[0] + v[1] * a[1] + v[2] * a[2]) / aa;
        const double r[3] = {v[0] - l * a[0], v[1] - l * a[1], v[2] - l * a[2]};
        if (l >= 0 && std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) <= 1e-12 * nv) best = std::min(best, l);
      }
      for (uint32_t k0 = 0; k0 < W; ++k0)
        for (uint32_t k1 = k0 + 1; k1 < W; ++k1) {
          const double* a = cj + 3 * k0;
          const double* b = cj + 3 * k1;
          double e[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
          double l[3];
          if (!solve3(a, b, e, v, l)) continue;
          if (l[0] >= 0 && l[1] >= 0 && std::abs(l[2]) * std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) <= 1e-12 * nv)
            best = std::min(best, l[0] + l[1]);
          for (uint32_t k2 = k1 + 1; k2 < W; ++k2) {
            if (!solve3(a, b, cj + 3 * k2, v, l)) continue;
            if (l[0] >= 0 && l[1] >= 0 && l[2] >= 0) best = std::min(best, l[0] + l[1] + l[2]);
          }
        }
      if (best == 1e300) return -1.0;
      worst = std::max(worst, best);
    }
  }
  return worst;
}

inline FastTables compute_fast_tables(const brille_grid_desc* d) {
  constexpr double two_pi = 6.283185307179586476925286766559;
  FastTables t;
  double Sp = 0, So = 0;
  face_affine_forms(d->prim_plane_a, d->prim_plane_b, d->prim_plane_c, d->n_faces, d->prim_cross_scale * two_pi,
                    t.prim_aff, &Sp);
  face_affine_forms(d->outer_plane_a, d->outer_plane_b, d->outer_plane_c, d->n_faces, d->outer_cross_scale * two_pi,
                    t.outer_aff, &So);
  // The exact face and wedge tests accept values within the zone tolerance's
  // band around zero (|x| <= T).  The cheap face values carry ~1e-15 S of
  // rounding error: a point is called inside when its smallest value exceeds
  // the margin, outside when it is below minus twice the margin, so the margin
  // must exceed both that error (1e-9 S leaves six orders) and T.  A
  // tolerance so loose that the band covers everything turns the fast tests off.
  double Trel = 0;
  const double T = zero_band(d->float_tolerance, d->approx_digits, &Trel);
  if (T >= 1e299) {
    t.aff_margin = 1e300;  // never decided fast
    return t;
  }
  t.aff_margin = std::max(std::max(1e-9 * std::max(Sp, So), 1e-11), 2.0 * T);
  if (d->has_wedge_seed && d->n_wedge && d->no_ir_mirroring) {
    // wedge seed orbit: u_j = R_j G* p0 with p0 scaled to unit Cartesian length
    const double* Gs = d->outer_metric_recip;
    const double* p0 = d->wedge_seed;
    double Gp[3];
    for (int i = 0; i < 3; ++i) Gp[i] = Gs[3 * i] * p0[0] + Gs[3 * i + 1] * p0[1] + Gs[3 * i + 2] * p0[2];
    double len = std::sqrt(p0[0] * Gp[0] + p0[1] * Gp[1] + p0[2] * Gp[2]);
    if (len > 0) {
      for (double& x : Gp) x /= len;
      t.orbit.resize(3 * (size_t)d->n_ops);
      for (uint32_t j = 0; j < d->n_ops; ++j) {
        const int32_t* R = d->rotations + 9 * j;
        for (int i = 0; i < 3; ++i) t.orbit[3 * j + i] = R[3 * i] * Gp[0] + R[3 * i + 1] * Gp[1] + R[3 * i + 2] * Gp[2];
      }
      // a lead beyond 1e-8 (scores are Cartesian lengths, |p0| = 1; their
      // rounding is ~1e-15) and beyond what the tolerant wedge test lets a
      // trailing operation pass with leaves the top scorer as the only
      // operation the exact search can accept
      const double lead = wedge_lead_bound(d, t.orbit);
      if (lead >= 0) {
        t.orbit_margin = std::max(1e-8, 2.0 * T * lead);
        t.has_orbit = true;
      } else {
        t.orbit.clear();
      }
    }
  }
  return t;
}

}  // namespace brille_amd
