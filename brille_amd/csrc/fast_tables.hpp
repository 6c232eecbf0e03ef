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

inline FastTables compute_fast_tables(const brille_grid_desc* d) {
  constexpr double two_pi = 6.283185307179586476925286766559;
  FastTables t;
  double Sp = 0, So = 0;
  face_affine_forms(d->prim_plane_a, d->prim_plane_b, d->prim_plane_c, d->n_faces, d->prim_cross_scale * two_pi,
                    t.prim_aff, &Sp);
  face_affine_forms(d->outer_plane_a, d->outer_plane_b, d->outer_plane_c, d->n_faces, d->outer_cross_scale * two_pi,
                    t.outer_aff, &So);
  // the cheap face values carry ~1e-15 S of rounding error; the exact ones
  // decide within ~1e-12 absolute: 1e-9 S leaves six orders of safety
  t.aff_margin = std::max(1e-9 * std::max(Sp, So), 1e-11);
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
      // scores are Cartesian lengths (|p0| = 1); the wedge test decides within ~1e-12
      t.orbit_margin = 1e-8;
      t.has_orbit = true;
    }
  }
  return t;
}

}  // namespace brille_amd
