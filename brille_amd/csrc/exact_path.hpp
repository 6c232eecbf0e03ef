// Stage (a) fold-in and stage (b) point location, restated exactly.
//
// These are the per-point routines of the plain kernel and the fallback of
// the fast kernel: the same loops as the reference, evaluated in FP64 with
// the reference's operation order.
#pragma once
#include "../../include/brille_amd.h"
#include "device_common.hpp"

namespace brille_amd {

// x - x is 0 exactly for every finite x and NaN for infinities and NaN
BRILLE_HD bool finite3(const double* Q) { return Q[0] - Q[0] == 0.0 && Q[1] - Q[1] == 0.0 && Q[2] - Q[2] == 0.0; }

// part_moveinto_prim for one point  src/bz_move.cpp:14-53
BRILLE_HD void part_moveinto_prim(const GridConst& g, const double* Q, double* q, int32_t* tau, Tol tbz) {
  const uint32_t F = g.F;
  int32_t last_shift[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) tau[i] = (int32_t)round(Q[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) q[i] = Q[i] - (double)tau[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) last_shift[i] = tau[i];
  uint32_t count = 0;
  while (count++ < F &&
         !inside_all_planes(F, g.prim_pa, g.prim_pb, g.prim_pc, q, g.prim_Gr, g.prim_Gd, g.prim_scale, tbz)) {
    // N_j = round(q.n_j / |tau_j|); the face with the largest positive N_j wins
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

BRILLE_HD void to_primitive(const GridConst& g, const double* Q, double* Qp) {
  mat_vec_i(Qp, g.P6t, Q);
#pragma unroll
  for (int i = 0; i < 3; ++i) Qp[i] /= 6.0;
}

BRILLE_HD void from_primitive(const GridConst& g, const double* qp, const int32_t* tp, double* q, int32_t* tau) {
  mat_vec_i(q, g.invPt, qp);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += g.invPt[3 * i + k] * tp[k];
    tau[i] = s;
  }
}

// BrillouinZone::moveinto for one point  src/bz_move.cpp:103-163
BRILLE_HD int moveinto_exact(const GridConst& g, const double* Q, double* q, int32_t* tau, Tol tbz) {
  if (!finite3(Q)) return BRILLE_PT_OUTSIDE_BZ;
  double Qp[3], qp[3];
  int32_t tp[3];
  if (g.transform_needed) to_primitive(g, Q, Qp);
  else { Qp[0] = Q[0]; Qp[1] = Q[1]; Qp[2] = Q[2]; }
  part_moveinto_prim(g, Qp, qp, tp, tbz);
  if (g.transform_needed) from_primitive(g, qp, tp, q, tau);
  else {
#pragma unroll
    for (int i = 0; i < 3; ++i) { q[i] = qp[i]; tau[i] = tp[i]; }
  }
  // isinside  src/bz.hpp:631-642
  if (!inside_all_planes(g.F, g.outer_pa, g.outer_pb, g.outer_pc, q, g.outer_Gr, g.outer_Gd, g.outer_scale, tbz))
    return BRILLE_PT_OUTSIDE_BZ;
  return BRILLE_PT_OK;
}

// BrillouinZone::_inside_wedge_outer  src/bz.hpp:757-763
BRILLE_HD bool inside_wedge(const GridConst& g, const double* p, Tol tbz, Tol tdef) {
  const uint32_t W = g.W;
  if (W == 0) return true;
  if (g.no_ir_mirroring) {
    for (uint32_t j = 0; j < W; ++j) {
      double n[3] = {g.wedge_n[3 * j], g.wedge_n[3 * j + 1], g.wedge_n[3 * j + 2]};
      if (!cmp_ge(metric_dot(n, p, g.outer_Gr), 0.0, tbz)) return false;
    }
    return true;
  }
  bool all_le = true;
  for (uint32_t j = 0; j < W && all_le; ++j) {
    double n[3] = {g.wedge_n[3 * j], g.wedge_n[3 * j + 1], g.wedge_n[3 * j + 2]};
    if (!cmp_le(metric_dot(n, p, g.outer_Gr), 0.0, tdef)) all_le = false;
  }
  if (all_le) return true;
  for (uint32_t j = 0; j < W; ++j) {
    double n[3] = {g.wedge_n[3 * j], g.wedge_n[3 * j + 1], g.wedge_n[3 * j + 2]};
    if (!cmp_ge(metric_dot(n, p, g.outer_Gr), 0.0, tdef)) return false;
  }
  return true;
}

// ordered wedge search  src/bz_move.cpp:262-285
BRILLE_HD int wedge_search_exact(const GridConst& g, double* q, uint32_t* rot, uint32_t* invrot, Tol tbz, Tol tdef) {
  if (inside_wedge(g, q, tbz, tdef)) {
    *rot = *invrot = g.identity;
    return BRILLE_PT_OK;
  }
  for (uint32_t j = 0; j < g.G; ++j) {
    double qj[3];
    mat_t_vec_i(qj, g.rot + 9 * j, q);
    if (inside_wedge(g, qj, tbz, tdef)) {
      q[0] = qj[0]; q[1] = qj[1]; q[2] = qj[2];
      *invrot = j;
      *rot = g.inv_index[j];
      return BRILLE_PT_OK;
    }
  }
  return BRILLE_PT_OUTSIDE_IR;
}

// find_bin / on_boundary  src/trellis_poly.hpp:66-81 (linear scan over the stored knots)
BRILLE_HD uint32_t find_bin_scan(const double* edges, uint32_t n, double x) {
  uint32_t d = 0;
  while (d < n && !(edges[d] > x)) ++d;
  if (d > n - 1 && x < edges[0]) d = 0;
  return d > 0 ? d - 1 : d;
}

// the same answer from O(1) arithmetic: guess floor((x - k0)/step) and move
// to the first stored knot > x, exactly as the scan defines it
BRILLE_HD uint32_t find_bin_fast(const double* edges, uint32_t n, double lo, double inv_step, double x) {
  if (!(x >= edges[0])) return find_bin_scan(edges, n, x);  // NaN or below the first knot
  double f = fmin((x - lo) * inv_step, (double)n);
  int64_t d = (int64_t)f + 1;  // candidate index of the first knot > x
  if (d < 1) d = 1;
  if (d > (int64_t)n) d = n;
  while (d > 0 && edges[d - 1] > x) --d;
  while (d < (int64_t)n && !(edges[d] > x)) ++d;
  return d > 0 ? (uint32_t)(d - 1) : 0u;
}

BRILLE_HD int on_boundary(const double* edges, uint32_t n, double x, uint32_t i, Tol tdef) {
  if (i + 2 < n && approx(edges[i + 1], x, tdef)) return 1;
  if (i > 0 && approx(edges[i], x, tdef)) return -1;
  return 0;
}

BRILLE_HD bool node_usable(const GridConst& g, const int64_t* sub, uint64_t* idx) {
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (sub[d] < 0 || sub[d] >= (int64_t)(g.nk[d] - 1)) return false;
  uint64_t s1 = g.nk[0] - 1, s2 = s1 * (g.nk[1] - 1);
  uint64_t k = (uint64_t)sub[0] + s1 * (uint64_t)sub[1] + s2 * (uint64_t)sub[2];
  *idx = k;
  uint8_t t = g.node_type[k];
  return t == BRILLE_NODE_CUBE || t == BRILLE_NODE_POLY;
}

// PolyTrellis::node_subscript  src/trellis_poly.hpp:382-434
template <bool FastBins>
BRILLE_HD bool node_subscript(const GridConst& g, const double* x, uint64_t* node, Tol tdef) {
  int64_t sub[3], ns[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    sub[d] = FastBins ? find_bin_fast(g.knots[d], g.nk[d], g.knot_lo[d], g.knot_inv_step[d], x[d])
                      : find_bin_scan(g.knots[d], g.nk[d], x[d]);
  uint64_t idx = 0;
  if (node_usable(g, sub, &idx)) { *node = idx; return true; }
  int close[3];
  int num_close = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    close[d] = on_boundary(g.knots[d], g.nk[d], x[d], (uint32_t)sub[d], tdef);
    num_close += close[d] != 0;
  }
  bool bad = true;
  if (num_close > 0)
    for (int i = 0; i < 3 && bad; ++i)
      if (close[i]) {
        ns[0] = sub[0]; ns[1] = sub[1]; ns[2] = sub[2];
        ns[i] += close[i];
        bad = !node_usable(g, ns, &idx);
      }
  if (bad && num_close > 1)
    for (int i = 0; i < 3 && bad; ++i)
      if (close[i])
        for (int j = 0; j < 3 && bad; ++j)
          if (close[j]) {
            ns[0] = sub[0]; ns[1] = sub[1]; ns[2] = sub[2];
            ns[i] += close[i];
            ns[j] += close[j];
            bad = !node_usable(g, ns, &idx);
          }
  if (bad && num_close > 2) {
#pragma unroll
    for (int i = 0; i < 3; ++i) ns[i] = sub[i] + close[i];
    bad = !node_usable(g, ns, &idx);
  }
  if (!bad) { *node = idx; return true; }
  return false;
}

// CubeNode::indices_weights  src/trellis_node.hpp:132-149, in fixed slots:
// slot i holds vertex 7-i and its weight.  A slot the reference drops (weight
// not > 0) keeps weight 0, so the non-zero slots in slot order are exactly the
// reference's (index, weight) list.
BRILLE_HD void cube_slots(const GridConst& g, uint32_t c, const double* x, uint32_t* idx, double* w, Tol tdef) {
  const uint32_t* vi = g.cube_v + 8 * (uint64_t)c;
  uint32_t v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = vi[i];
  const double* v0 = g.verts + 3 * (uint64_t)v[0];
  const double* v7 = g.verts + 3 * (uint64_t)v[7];
  double node_volume = 1.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) node_volume *= fabs(v0[k] - v7[k]);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double* vv = g.verts + 3 * (uint64_t)v[i];
    double p = 1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) p *= fabs(x[k] - vv[k]);
    double wi = p / node_volume;
    idx[i] = v[7 - i];
    w[i] = cmp_gt(wi, 0.0, tdef) ? wi : 0.0;
  }
}

// PolyNode::tetrahedra_might_contain  src/trellis_node.hpp:349-364
BRILLE_HD double tet_might_contain(const GridConst& g, uint64_t t, const double* x, Tol tdef) {
  const double* ci = g.tet_c + 4 * t;
  double v0 = ci[0] - x[0], v1 = ci[1] - x[1], v2 = ci[2] - x[2];
  double d2 = 0.0, r2 = ci[3] * ci[3];
  d2 += v0 * v0;
  d2 += v1 * v1;
  d2 += v2 * v2;
  return (d2 < r2 || approx(d2, r2, tdef)) ? 0. : -d2;
}

// PolyNode::tetrahedra_contains  src/trellis_node.hpp:319-339
BRILLE_HD double tet_contains(const GridConst& g, uint64_t t, const double* x, double* w, bool shortcut, Tol tdef) {
  if (shortcut) {
    double away = tet_might_contain(g, t, x, tdef);
    if (away < 0.) return away;
  }
  const uint32_t* vt = g.tet_v + 4 * t;
  const double* p0 = g.verts + 3 * (uint64_t)vt[0];
  const double* p1 = g.verts + 3 * (uint64_t)vt[1];
  const double* p2 = g.verts + 3 * (uint64_t)vt[2];
  const double* p3 = g.verts + 3 * (uint64_t)vt[3];
  double a0[3] = {p0[0], p0[1], p0[2]}, a1[3] = {p1[0], p1[1], p1[2]};
  double a2[3] = {p2[0], p2[1], p2[2]}, a3[3] = {p3[0], p3[1], p3[2]};
  double vol6 = g.tet_vol[t] * 6.0;
  w[0] = orient3d_xyz(x, a1, a2, a3) / vol6;
  w[1] = orient3d_xyz(a0, x, a2, a3) / vol6;
  w[2] = orient3d_xyz(a0, a1, x, a3) / vol6;
  w[3] = orient3d_xyz(a0, a1, a2, x) / vol6;
  bool any_neg = false;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (w[j] < 0. && !approx(w[j], 0., tdef)) any_neg = true;
  if (any_neg) {
    double m = w[0];
#pragma unroll
    for (int j = 1; j < 4; ++j)
      if (w[j] < m) m = w[j];
    return m;
  }
  return 0.;
}

// slot j of tetrahedron t; weight 0 where the reference drops the vertex
// (approximately zero weight, src/trellis_node.hpp:300-305)
BRILLE_HD void tet_slots(const GridConst& g, uint64_t t, const double* ww, uint32_t* idx, double* w, Tol tdef) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    idx[j] = g.tet_v[4 * t + j];
    w[j] = approx(ww[j], 0., tdef) ? 0.0 : ww[j];
  }
}

// PolyNode::indices_weights, should_contain=true  src/trellis_node.hpp:275-308.
// Fills slots 0..3 and returns the tetrahedron used (kNoTet: node has none).
BRILLE_HD uint32_t poly_slots(const GridConst& g, uint32_t p, const double* x, uint32_t* idx, double* w, Tol tdef) {
  const uint32_t t0 = g.poly_off[p], t1 = g.poly_off[p + 1];
  double ww[4];
  double best = 0.0;
  uint32_t best_t = t0;
  for (uint32_t t = t0; t < t1; ++t) {
    double mn = tet_contains(g, t, x, ww, true, tdef);
    if (mn >= 0.) {
      tet_slots(g, t, ww, idx, w, tdef);
      return t;
    }
    if (t == t0 || mn > best) { best = mn; best_t = t; }
  }
  if (t1 == t0) return kNoTet;
  tet_contains(g, best_t, x, ww, false, tdef);
  tet_slots(g, best_t, ww, idx, w, tdef);
  return best_t;
}

}  // namespace brille_amd
