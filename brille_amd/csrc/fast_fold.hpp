// Fast stage (a): the exact fold-in's decisions, taken on cheap bounds first.
//
// The result of the fold-in is decided by signs and tolerance compares only:
// "is q inside every zone face", "which lattice vector to subtract", "which
// point-group operation puts q in the wedge".  The q and tau it produces are
// always computed with the exact reference arithmetic.  Here each decision is
// first taken on an algebraically identical but cheaper expression, whose
// rounding error is far below a safety margin; only when the cheap value lies
// within that margin of the decision boundary is the exact expression
// evaluated.  The outcome is therefore identical to moveinto_exact /
// wedge_search_exact, bit for bit.
//
//  * zone faces: pseudo_orient3d(a, b, c, x) equals det[a-x, b-x, c-x] in
//    Cartesian space, i.e. alpha - beta.x with alpha = det[a, b, c] and
//    beta = B^T (a x b + b x c + c x a), precomputed per face (prim_aff, outer_aff);
//  * wedge: the wedge is the Dirichlet domain of a seed p0, so R_j^T q lies in it
//    exactly when score_j = q . (R_j G* p0) is the largest of all scores
//    (orbit); a winner ahead of the runner-up by more than the margin is the only
//    operation whose image passes the tolerant wedge test.
#pragma once
#include "exact_path.hpp"

namespace brille_amd {

// +1 inside, -1 outside, 0 undecided; m is the smallest affine face value
BRILLE_HD int fast_inside(uint32_t F, const double* aff, const double* x, double margin) {
  double m = 1e300;
  for (uint32_t j = 0; j < F; ++j) {
    const double* a = aff + 4 * j;
    double o = fma(-a[0], x[0], fma(-a[1], x[1], fma(-a[2], x[2], a[3])));
    m = fmin(m, o);
  }
  if (m > margin) return 1;
  if (m < -2.0 * margin) return -1;
  return 0;
}

BRILLE_HD bool inside_prim(const GridConst& g, const double* q, Tol tbz) {
  int f = fast_inside(g.F, g.prim_aff, q, g.aff_margin);
  if (f) return f > 0;
  return inside_all_planes(g.F, g.prim_pa, g.prim_pb, g.prim_pc, q, g.prim_Gr, g.prim_Gd, g.prim_scale, tbz);
}

BRILLE_HD bool inside_outer(const GridConst& g, const double* q, Tol tbz) {
  int f = fast_inside(g.F, g.outer_aff, q, g.aff_margin);
  if (f) return f > 0;
  return inside_all_planes(g.F, g.outer_pa, g.outer_pb, g.outer_pc, q, g.outer_Gr, g.outer_Gd, g.outer_scale, tbz);
}

// part_moveinto_prim with fast inside tests  src/bz_move.cpp:14-53
BRILLE_HD void part_moveinto_prim_fast(const GridConst& g, const double* Q, double* q, int32_t* tau, Tol tbz) {
  const uint32_t F = g.F;
  int32_t last_shift[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) tau[i] = (int32_t)round(Q[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) q[i] = Q[i] - (double)tau[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) last_shift[i] = tau[i];
  uint32_t count = 0;
  while (count++ < F && !inside_prim(g, q, tbz)) {
    int32_t max_nm = 0;
    uint32_t max_at = 0;
    double dmax = 0.0;
    for (uint32_t j = 0; j < F; ++j) {
      double n[3] = {g.prim_n[3 * j], g.prim_n[3 * j + 1], g.prim_n[3 * j + 2]};
      double dj = metric_dot(q, n, g.prim_Gr);
      int32_t Nj = (int32_t)round(dj / g.prim_tau_lens[j]);
      const int32_t* tj = g.prim_taus + 3 * j;
      bool nonzero = (tj[0] + last_shift[0]) != 0 || (tj[1] + last_shift[1]) != 0 || (tj[2] + last_shift[2]) != 0;
      if (Nj > 0 && Nj >= max_nm && (0 == max_nm || (nonzero && dj > dmax))) {
        max_at = j;
        max_nm = Nj;
        dmax = dj;
      }
    }
    if (max_nm > 0) {
      const int32_t* tm = g.prim_taus + 3 * max_at;
#pragma unroll
      for (int i = 0; i < 3; ++i) q[i] -= (double)tm[i] * (double)max_nm;
#pragma unroll
      for (int i = 0; i < 3; ++i) tau[i] += tm[i] * max_nm;
#pragma unroll
      for (int i = 0; i < 3; ++i) last_shift[i] = tm[i] * max_nm;
    }
  }
}

BRILLE_HD int fast_moveinto(const GridConst& g, const double* Q, double* q, int32_t* tau, Tol tbz) {
  if (!finite3(Q)) return BRILLE_PT_OUTSIDE_BZ;
  double Qp[3], qp[3];
  int32_t tp[3];
  if (g.transform_needed) to_primitive(g, Q, Qp);
  else { Qp[0] = Q[0]; Qp[1] = Q[1]; Qp[2] = Q[2]; }
  part_moveinto_prim_fast(g, Qp, qp, tp, tbz);
  if (g.transform_needed) from_primitive(g, qp, tp, q, tau);
  else {
#pragma unroll
    for (int i = 0; i < 3; ++i) { q[i] = qp[i]; tau[i] = tp[i]; }
  }
  if (!inside_outer(g, q, tbz)) return BRILLE_PT_OUTSIDE_BZ;
  return BRILLE_PT_OK;
}

BRILLE_HD int fast_wedge_search(const GridConst& g, double* q, uint32_t* rot, uint32_t* invrot, Tol tbz, Tol tdef) {
  if (g.W == 0 || !g.has_seed || !g.no_ir_mirroring) return wedge_search_exact(g, q, rot, invrot, tbz, tdef);
  double best = -1e300, second = -1e300;
  uint32_t jb = 0;
  for (uint32_t j = 0; j < g.G; ++j) {
    const double* u = g.orbit + 3 * j;
    double s = fma(q[0], u[0], fma(q[1], u[1], q[2] * u[2]));
    if (s > best) { second = best; best = s; jb = j; }
    else if (s > second) second = s;
  }
  if (best - second > g.orbit_margin) {
    if (jb == g.identity) {
      *rot = *invrot = g.identity;
      return BRILLE_PT_OK;
    }
    double qj[3];
    mat_t_vec_i(qj, g.rot + 9 * jb, q);
    q[0] = qj[0]; q[1] = qj[1]; q[2] = qj[2];
    *invrot = jb;
    *rot = g.inv_index[jb];
    return BRILLE_PT_OK;
  }
  return wedge_search_exact(g, q, rot, invrot, tbz, tdef);
}

}  // namespace brille_amd
