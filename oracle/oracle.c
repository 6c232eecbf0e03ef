/*
 * oracle.c -- CPU restatement of brille's ir_interpolate_at path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU kernels are
 * compared against (tests/, __graft_entry__.smoke(), and the cpu_baseline leg
 * of bench.py).  The product path never links, loads or calls it.
 *
 * It follows the reference loops literally, operation by operation, so that
 * every floating-point decision (rounding, tolerance compares, orientation
 * signs) is taken on the same IEEE binary64 values as the reference computes
 * when built without FMA contraction.  Each function cites the reference
 * lines it restates (paths relative to the brille source tree).
 *
 * Parallel structure mirrors the reference: stage (a) serial
 * (src/bz_move.cpp:135-141, 262-285), stages (b)+(c1) `omp parallel for
 * schedule(dynamic)` (src/trellis_poly.hpp:330), stage (c2) `omp parallel for
 * schedule(static)` (src/interpolator_gamma.tpp:79-81).
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/brille_amd.h"

#define TWO_PI 6.283185307179586476925286766559

/* ------------------------------------------------------------------------ */
/* approx_float::tols / _scalar  src/approx_float.hpp:79-100, 171-177       */
/* ------------------------------------------------------------------------ */
typedef struct { double rel, abs; } tol_t;

static tol_t make_tols(double Ttol, int tol) {
  tol_t t;
  double Trel = DBL_EPSILON;
  double Tabs = 5.0 / 1000000000000000.0;
  Trel *= (double)tol * 10000.0; /* static_cast<T>(tol)*static_cast<T>(TOL_MULT) */
  if (Ttol > Trel) Trel = Ttol;
  if (Ttol > Tabs) Tabs = Ttol;
  t.rel = Trel;
  t.abs = Tabs;
  return t;
}

static int approx(double a, double b, tol_t t) {
  double x = fabs(a - b);
  return x <= t.abs + t.rel * fabs(a + b) || x < DBL_MIN;
}

/* Comparer cmp::ge / cmp::le / cmp::gt  src/comparisons.hpp:72-86 */
static int cmp_ge(double a, double b, tol_t t) { return approx(a, b, t) || a > b; }
static int cmp_le(double a, double b, tol_t t) { return approx(a, b, t) || a < b; }
static int cmp_gt(double a, double b, tol_t t) { return !approx(a, b, t) && a > b; }

/* ------------------------------------------------------------------------ */
/* small linear algebra, in the reference's accumulation order               */
/* ------------------------------------------------------------------------ */
/* utils::multiply_matrix_vector / mul_mat_vec  src/utilities.tpp:44-55 */
static void mat_vec_d(double* c, const double* A, const double* b) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += A[3 * i + k] * b[k];
    c[i] = s;
  }
}
static void mat_vec_i(double* c, const int32_t* A, const double* b) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += (double)A[3 * i + k] * b[k];
    c[i] = s;
  }
}
static void mat_vec_ii(int32_t* c, const int32_t* A, const int32_t* b) {
  for (int i = 0; i < 3; ++i) {
    int32_t s = 0;
    for (int k = 0; k < 3; ++k) s += A[3 * i + k] * b[k];
    c[i] = s;
  }
}

/* same_lattice_dot(x, y, metric)  src/array_functions.hpp:238-246 */
static double metric_dot(const double* x, const double* y, const double* M) {
  double tmp[3];
  mat_vec_d(tmp, M, x);
  double out = 0.0;
  for (int i = 0; i < 3; ++i) out += tmp[i] * y[i];
  return out;
}

/* vector_cross  src/utilities.hpp:369-376 */
static void cross3(double* c, const double* a, const double* b) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/*
 * pseudo_orient3d for lattice vectors: dot(a-d, cross(b-d, c-d)) with the
 * LVec cross (src/array_functions.hpp:190-202: plain cross of the components,
 * scaled by V/(2pi), then LVec::star src/array_lvec_methods.tpp:53-66:
 * real metric times the vector divided by 2pi) and the same-lattice metric
 * dot (src/array_functions.hpp:249-291).  src/geometry.hpp:118-138
 */
static double orient3d_lattice(const double* a, const double* b, const double* c,
                               const double* d, const double* Grecip,
                               const double* Greal, double scale) {
  double u[3], v[3], w[3], cr[3], st[3];
  for (int i = 0; i < 3; ++i) { u[i] = a[i] - d[i]; v[i] = b[i] - d[i]; w[i] = c[i] - d[i]; }
  cross3(cr, v, w);
  for (int i = 0; i < 3; ++i) cr[i] *= scale;
  mat_vec_d(st, Greal, cr);
  for (int i = 0; i < 3; ++i) st[i] /= TWO_PI;
  return metric_dot(u, st, Grecip);
}

/* pseudo_orient3d for bare Cartesian arrays  src/geometry.hpp:118-138 */
static double orient3d_xyz(const double* a, const double* b, const double* c, const double* d) {
  double u[3], v[3], w[3], cr[3];
  for (int i = 0; i < 3; ++i) { u[i] = a[i] - d[i]; v[i] = b[i] - d[i]; w[i] = c[i] - d[i]; }
  cross3(cr, v, w);
  double out = 0.0;
  for (int i = 0; i < 3; ++i) out += u[i] * cr[i];
  return out;
}

/* point_inside_all_planes  src/geometry.hpp:413-419 */
static int inside_all_planes(uint32_t F, const double* pa, const double* pb, const double* pc,
                             const double* x, const double* Grecip, const double* Greal,
                             double scale, tol_t t) {
  double vmin = 0.0;
  for (uint32_t j = 0; j < F; ++j) {
    double o = orient3d_lattice(pa + 3 * j, pb + 3 * j, pc + 3 * j, x, Grecip, Greal, scale);
    if (j == 0 || o < vmin) vmin = o; /* std::min_element keeps the first minimum */
  }
  return vmin > 0 || approx(vmin, 0.0, t);
}

/* ------------------------------------------------------------------------ */
/* stage (a): moveinto + wedge search                                        */
/* ------------------------------------------------------------------------ */
/* part_moveinto_prim, one point  src/bz_move.cpp:14-53 */
static void part_moveinto_prim(const brille_grid_desc* g, const double* Q, double* q_out,
                               int32_t* tau_out, tol_t tbz) {
  const uint32_t F = g->n_faces;
  int32_t tau[3], last_shift[3];
  double q[3];
  for (int i = 0; i < 3; ++i) tau[i] = (int32_t)round(Q[i]);
  for (int i = 0; i < 3; ++i) q[i] = Q[i] - (double)tau[i];
  for (int i = 0; i < 3; ++i) last_shift[i] = tau[i];
  uint32_t count = 0;
  double dots[64];
  int32_t N[64];
  while (count++ < F && !inside_all_planes(F, g->prim_plane_a, g->prim_plane_b, g->prim_plane_c, q,
                                          g->prim_metric_recip, g->prim_metric_real,
                                          g->prim_cross_scale, tbz)) {
    int any = 0;
    for (uint32_t j = 0; j < F; ++j) {
      dots[j] = metric_dot(q, g->prim_normals + 3 * j, g->prim_metric_recip);
      N[j] = (int32_t)round(dots[j] / g->prim_tau_lens[j]);
      if (N[j] > 0) any = 1;
    }
    if (any) {
      int32_t max_nm = 0;
      uint32_t max_at = 0;
      for (uint32_t j = 0; j < F; ++j) {
        const int32_t* tj = g->prim_taus + 3 * j;
        /* norm(taus_j + last_shift) > 0: the metric is positive definite, so
         * this is "the integer vector is not zero" */
        int nonzero = (tj[0] + last_shift[0]) != 0 || (tj[1] + last_shift[1]) != 0 ||
                      (tj[2] + last_shift[2]) != 0;
        if (N[j] > 0 && N[j] >= max_nm &&
            (0 == max_nm || (nonzero && dots[j] > dots[max_at]))) {
          max_at = j;
          max_nm = N[j];
        }
      }
      const int32_t* tm = g->prim_taus + 3 * max_at;
      for (int i = 0; i < 3; ++i) q[i] -= (double)tm[i] * (double)max_nm;
      for (int i = 0; i < 3; ++i) tau[i] += tm[i] * max_nm;
      for (int i = 0; i < 3; ++i) last_shift[i] = tm[i] * max_nm;
    }
  }
  for (int i = 0; i < 3; ++i) { q_out[i] = q[i]; tau_out[i] = tau[i]; }
}

/* BrillouinZone::moveinto, one point  src/bz_move.cpp:103-163 */
static int moveinto(const brille_grid_desc* g, const double* Q, double* q, int32_t* tau, tol_t tbz) {
  double Qp[3], qp[3];
  int32_t taup[3];
  if (!(isfinite(Q[0]) && isfinite(Q[1]) && isfinite(Q[2]))) return BRILLE_PT_OUTSIDE_BZ;
  if (g->transform_needed) {
    /* parallel_transform_to_primitive  src/transform.hpp:161-199 */
    mat_vec_i(Qp, g->to_primitive_6Pt, Q);
    for (int i = 0; i < 3; ++i) Qp[i] /= 6.0;
  } else {
    for (int i = 0; i < 3; ++i) Qp[i] = Q[i];
  }
  part_moveinto_prim(g, Qp, qp, taup, tbz);
  if (g->transform_needed) {
    /* parallel_transform_from_primitive  src/transform.hpp:201-239 */
    mat_vec_i(q, g->from_primitive_invPt, qp);
    mat_vec_ii(tau, g->from_primitive_invPt, taup);
  } else {
    for (int i = 0; i < 3; ++i) { q[i] = qp[i]; tau[i] = taup[i]; }
  }
  /* isinside -> _first.contains(q, float_tolerance, approx_tolerance)
   * src/bz.hpp:631-642, src/polyhedron_faces.hpp:316-337 */
  if (!inside_all_planes(g->n_faces, g->outer_plane_a, g->outer_plane_b, g->outer_plane_c, q,
                         g->outer_metric_recip, g->outer_metric_real, g->outer_cross_scale, tbz))
    return BRILLE_PT_OUTSIDE_BZ;
  return BRILLE_PT_OK;
}

/* BrillouinZone::_inside_wedge_outer  src/bz.hpp:757-763, Array2::all src/array2.tpp:665-672 */
static int inside_wedge(const brille_grid_desc* g, const double* p, tol_t tbz, tol_t tdef) {
  const uint32_t W = g->n_wedge;
  if (W == 0) return 1;
  if (g->no_ir_mirroring) {
    for (uint32_t j = 0; j < W; ++j)
      if (!cmp_ge(metric_dot(g->wedge_normals + 3 * j, p, g->outer_metric_recip), 0.0, tbz)) return 0;
    return 1;
  }
  /* cmp::le_ge: all(le) || all(ge), with the default tolerance */
  int all_le = 1, all_ge = 1;
  for (uint32_t j = 0; j < W && all_le; ++j)
    if (!cmp_le(metric_dot(g->wedge_normals + 3 * j, p, g->outer_metric_recip), 0.0, tdef)) all_le = 0;
  if (all_le) return 1;
  for (uint32_t j = 0; j < W && all_ge; ++j)
    if (!cmp_ge(metric_dot(g->wedge_normals + 3 * j, p, g->outer_metric_recip), 0.0, tdef)) all_ge = 0;
  return all_ge;
}

/* BrillouinZone::ir_moveinto, one point  src/bz_move.cpp:182, 257-296 */
static int ir_moveinto(const brille_grid_desc* g, const double* Q, double* q, int32_t* tau,
                       uint32_t* rot, uint32_t* invrot, tol_t tbz, tol_t tdef) {
  int st = moveinto(g, Q, q, tau, tbz);
  if (st != BRILLE_PT_OK) return st;
  if (inside_wedge(g, q, tbz, tdef)) {
    *rot = *invrot = g->identity_index;
    return BRILLE_PT_OK;
  }
  for (uint32_t j = 0; j < g->n_ops; ++j) {
    const int32_t* R = g->rotations + 9 * j;
    int32_t Rt[9];
    double qj[3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Rt[3 * a + b] = R[3 * b + a];
    mat_vec_i(qj, Rt, q);
    if (inside_wedge(g, qj, tbz, tdef)) {
      for (int i = 0; i < 3; ++i) q[i] = qj[i];
      *invrot = j;
      *rot = g->inverse_index[j];
      return BRILLE_PT_OK;
    }
  }
  return BRILLE_PT_OUTSIDE_IR;
}

/* ------------------------------------------------------------------------ */
/* stage (b): point location                                                 */
/* ------------------------------------------------------------------------ */
/* find_bin / on_boundary  src/trellis_poly.hpp:66-81 */
static uint32_t find_bin(const double* edges, uint32_t n, double x) {
  uint32_t d = 0;
  while (d < n && !(edges[d] > x)) ++d;
  if (d > n - 1 && x < edges[0]) d = 0;
  return d > 0 ? d - 1 : d;
}
static int on_boundary(const double* edges, uint32_t n, double x, uint32_t i, tol_t tdef) {
  if (i + 2 < n && approx(edges[i + 1], x, tdef)) return 1;
  if (i > 0 && approx(edges[i], x, tdef)) return -1;
  return 0;
}

static int node_is_null(uint8_t t) { return t != BRILLE_NODE_CUBE && t != BRILLE_NODE_POLY; }

static int sub_ok_not_null(const brille_grid_desc* g, const int64_t* sub, uint64_t* idx) {
  for (int d = 0; d < 3; ++d)
    if (sub[d] < 0 || (uint64_t)sub[d] >= (uint64_t)(g->n_knots[d] - 1)) return 0;
  uint64_t s1 = g->n_knots[0] - 1, s2 = s1 * (g->n_knots[1] - 1);
  uint64_t k = (uint64_t)sub[0] + s1 * (uint64_t)sub[1] + s2 * (uint64_t)sub[2];
  *idx = k;
  return !node_is_null(g->node_type[k]);
}

/* PolyTrellis::node_subscript  src/trellis_poly.hpp:382-434
 * returns 1 with *node set when a usable node was found */
static int node_subscript(const brille_grid_desc* g, const double* x, uint64_t* node) {
  int64_t sub[3], newsub[3];
  uint64_t idx = 0;
  for (int d = 0; d < 3; ++d) sub[d] = find_bin(g->knots[d], g->n_knots[d], x[d]);
  if (sub_ok_not_null(g, sub, &idx)) { *node = idx; return 1; }
  tol_t tdef = make_tols(0.0, 1);
  int close[3];
  int num_close = 0;
  for (int d = 0; d < 3; ++d) {
    close[d] = on_boundary(g->knots[d], g->n_knots[d], x[d], (uint32_t)sub[d], tdef);
    if (close[d]) ++num_close;
  }
  int bad = 1;
  if (num_close > 0)
    for (int i = 0; i < 3 && bad; ++i)
      if (close[i]) {
        memcpy(newsub, sub, sizeof(sub));
        newsub[i] += close[i];
        bad = !sub_ok_not_null(g, newsub, &idx);
      }
  if (bad && num_close > 1)
    for (int i = 0; i < 3 && bad; ++i)
      if (close[i])
        for (int j = 0; j < 3 && bad; ++j)
          if (close[j]) {
            memcpy(newsub, sub, sizeof(sub));
            newsub[i] += close[i];
            newsub[j] += close[j];
            bad = !sub_ok_not_null(g, newsub, &idx);
          }
  if (bad && num_close > 2) {
    for (int i = 0; i < 3; ++i) newsub[i] = sub[i] + close[i];
    bad = !sub_ok_not_null(g, newsub, &idx);
  }
  if (!bad) { *node = idx; return 1; }
  return 0;
}

/* CubeNode::indices_weights  src/trellis_node.hpp:132-149 */
static int cube_weights(const brille_grid_desc* g, uint64_t c, const double* x, uint32_t* idx,
                        double* w, tol_t tdef) {
  const uint32_t* vi = g->cube_vertices + 8 * c;
  const double* v0 = g->vertices + 3 * (uint64_t)vi[0];
  const double* v7 = g->vertices + 3 * (uint64_t)vi[7];
  double node_volume = 1.0;
  for (int k = 0; k < 3; ++k) node_volume *= fabs(v0[k] - v7[k]);
  int n = 0;
  for (int i = 0; i < 8; ++i) {
    const double* vv = g->vertices + 3 * (uint64_t)vi[i];
    double p = 1.0;
    for (int k = 0; k < 3; ++k) p *= fabs(x[k] - vv[k]);
    double wi = p / node_volume;
    if (cmp_gt(wi, 0.0, tdef)) {
      idx[n] = vi[7 - i];
      w[n] = wi;
      ++n;
    }
  }
  return n;
}

/* PolyNode::tetrahedra_might_contain  src/trellis_node.hpp:349-364 */
static double tet_might_contain(const brille_grid_desc* g, uint64_t t, const double* x, tol_t tdef) {
  const double* ci = g->tet_circumsphere + 4 * t;
  double v[3];
  v[0] = ci[0] - x[0];
  v[1] = ci[1] - x[1];
  v[2] = ci[2] - x[2];
  double d2 = 0.0, r2 = ci[3] * ci[3];
  for (int i = 0; i < 3; ++i) d2 += v[i] * v[i];
  return (d2 < r2 || approx(d2, r2, tdef)) ? 0. : -d2;
}

/* PolyNode::tetrahedra_contains  src/trellis_node.hpp:319-339 */
static double tet_contains(const brille_grid_desc* g, uint64_t t, const double* x, double* w,
                           int allow_shortcut, tol_t tdef) {
  if (allow_shortcut) {
    double away = tet_might_contain(g, t, x, tdef);
    if (away < 0.) return away;
  }
  const uint32_t* vt = g->tet_vertices + 4 * t;
  const double* v0 = g->vertices + 3 * (uint64_t)vt[0];
  const double* v1 = g->vertices + 3 * (uint64_t)vt[1];
  const double* v2 = g->vertices + 3 * (uint64_t)vt[2];
  const double* v3 = g->vertices + 3 * (uint64_t)vt[3];
  double vol6 = g->tet_volume[t] * 6.0;
  w[0] = orient3d_xyz(x, v1, v2, v3) / vol6;
  w[1] = orient3d_xyz(v0, x, v2, v3) / vol6;
  w[2] = orient3d_xyz(v0, v1, x, v3) / vol6;
  w[3] = orient3d_xyz(v0, v1, v2, x) / vol6;
  int any_neg = 0;
  for (int j = 0; j < 4; ++j)
    if (w[j] < 0. && !approx(w[j], 0., tdef)) any_neg = 1;
  if (any_neg) {
    double m = w[0];
    for (int j = 1; j < 4; ++j)
      if (w[j] < m) m = w[j];
    return m;
  }
  return 0.;
}

/* PolyNode::indices_weights with should_contain=true  src/trellis_node.hpp:275-308 */
static int poly_weights(const brille_grid_desc* g, uint64_t p, const double* x, uint32_t* idx,
                        double* w, uint32_t* tet_out, tol_t tdef) {
  uint32_t t0 = g->poly_tet_offset[p], t1 = g->poly_tet_offset[p + 1];
  double ww[4];
  double best = 0.0;
  uint32_t best_t = t0;
  for (uint32_t t = t0; t < t1; ++t) {
    double mn = tet_contains(g, t, x, ww, 1, tdef);
    if (mn >= 0.) {
      int n = 0;
      for (int j = 0; j < 4; ++j)
        if (!approx(ww[j], 0., tdef)) { idx[n] = g->tet_vertices[4 * (uint64_t)t + j]; w[n] = ww[j]; ++n; }
      *tet_out = t;
      return n;
    }
    if (t == t0 || mn > best) { best = mn; best_t = t; } /* std::max_element: first maximum */
  }
  if (t1 == t0) return 0;
  tet_contains(g, best_t, x, ww, 0, tdef);
  int n = 0;
  for (int j = 0; j < 4; ++j)
    if (!approx(ww[j], 0., tdef)) { idx[n] = g->tet_vertices[4 * (uint64_t)best_t + j]; w[n] = ww[j]; ++n; }
  *tet_out = best_t;
  return n;
}

/* ------------------------------------------------------------------------ */
/* stage (c1): weighted sums  src/interpolator_at.tpp:92-127                  */
/* ------------------------------------------------------------------------ */
/* PermutationTable::safe_get  src/permutation_table.hpp:193-198, 214 */
static uint32_t perm_at(const brille_perm_desc* pd, uint64_t V, uint32_t B, uint32_t pivot,
                        uint32_t idx, uint32_t b) {
  if (!pd || pd->n_keys == 0 || pivot == idx) return b;
  uint64_t key = (uint64_t)pivot * V + idx;
  uint64_t lo = 0, hi = pd->n_keys;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (pd->keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  if (lo < pd->n_keys && pd->keys[lo] == key)
    return pd->perms[(uint64_t)pd->perm_index[lo] * B + b];
  return b;
}

/* utils::antiphase  src/utilities.tpp:567-577 */
static void antiphase(uint32_t n, const double* a, const double* b, double* re, double* im) {
  double real_dot = 0, imag_dot = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
    real_dot += ar * br + ai * bi;
    imag_dot += ar * bi - ai * br;
  }
  double th = -1.0 * atan2(imag_dot, real_dot);
  *re = 1.0 * cos(th);
  *im = 1.0 * sin(th);
}

/* ------------------------------------------------------------------------ */
/* stage (c2): rotations                                                     */
/* ------------------------------------------------------------------------ */
static int is_identity(const int32_t* R) {
  static const int32_t e[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int i = 0; i < 9; ++i) if (R[i] != e[i]) return 0;
  return 1;
}

/* complex (re,im) pairs times an integer or real 3x3, mul_arrays src/utilities.tpp:58-63 */
static void mat_cvec_i(double* c, const int32_t* A, const double* v) {
  for (int i = 0; i < 3; ++i) {
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < 3; ++k) { sr += (double)A[3 * i + k] * v[2 * k]; si += (double)A[3 * i + k] * v[2 * k + 1]; }
    c[2 * i] = sr; c[2 * i + 1] = si;
  }
}
static void mat_cvec_d(double* c, const double* A, const double* v) {
  for (int i = 0; i < 3; ++i) {
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < 3; ++k) { sr += A[3 * i + k] * v[2 * k]; si += A[3 * i + k] * v[2 * k + 1]; }
    c[2 * i] = sr; c[2 * i + 1] = si;
  }
}
static void mat_vec_real_i(double* c, const int32_t* A, const double* v) { mat_vec_i(c, A, v); }

/* 3x3 matrix products on real or complex element storage, mul_mat_mat src/utilities.hpp:119 */
static void matmat_x_i(double* C, const double* X, const int32_t* R, int cplx) {
  /* C = X * R, X stored with stride (1 or 2 doubles per element) */
  int s = cplx ? 2 : 1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int p = 0; p < s; ++p) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += X[s * (3 * i + k) + p] * (double)R[3 * k + j];
        C[s * (3 * i + j) + p] = acc;
      }
}
static void matmat_i_x(double* C, const int32_t* R, const double* X, int cplx) {
  int s = cplx ? 2 : 1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int p = 0; p < s; ++p) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += (double)R[3 * i + k] * X[s * (3 * k + j) + p];
        C[s * (3 * i + j) + p] = acc;
      }
}
static void transpose_i(int32_t* Rt, const int32_t* R) {
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) Rt[3 * a + b] = R[3 * b + a];
}

/*
 * Interpolator::rotate_in_place for one point.  src/interpolator.hpp:387-428;
 * rip_real src/interpolator_real.tpp:19-64; rip_recip src/interpolator_recip.tpp:19-57;
 * rip_axial src/interpolator_axial.tpp:19-60; rip_gamma_complex
 * src/interpolator_gamma.tpp:49-138 (vector part; tensors handled as rip_real)
 */
static int rotate_point(const brille_grid_desc* g, const brille_data_desc* d,
                        const brille_gamma_desc* gt, double* x, const double* irq,
                        uint32_t r, uint32_t ir, double* tA) {
  const uint32_t B = d->branches;
  const uint32_t s_ = d->elements[0] + d->elements[1] + d->elements[2];
  const uint32_t nv = d->elements[1] / 3u, nm = d->elements[2] / 9u, ns = d->elements[0];
  const int cplx = d->is_complex;
  const int st = cplx ? 2 : 1;
  if (nv == 0 && nm == 0) return BRILLE_OK;
  const int32_t* R = g->rotations + 9 * (uint64_t)r;
  const int32_t* iR = g->rotations + 9 * (uint64_t)ir;
  int32_t Rt[9], iRt[9];
  double tmpv[6], tmpm[18];
  switch (d->length_unit) {
    case BRILLE_LU_REAL_LATTICE:
      if (d->rotates_like == BRILLE_ROT_GAMMA) break;
      if (d->rotates_like == BRILLE_ROT_VECTOR || d->rotates_like == BRILLE_ROT_PSEUDOVECTOR) {
        if (is_identity(R)) return BRILLE_OK;
        int axial = d->rotates_like == BRILLE_ROT_PSEUDOVECTOR;
        double detR = 0.0;
        if (axial) {
          detR = (double)(R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                          R[2] * (R[3] * R[7] - R[4] * R[6]));
        }
        for (uint32_t b = 0; b < B; ++b) {
          uint32_t o = b * s_ + ns;
          for (uint32_t v = 0; v < nv; ++v) {
            double* xv = x + st * o;
            if (cplx) mat_cvec_i(tmpv, axial ? iR : R, xv); else mat_vec_real_i(tmpv, axial ? iR : R, xv);
            for (int j = 0; j < 3 * st; ++j) xv[j] = axial ? detR * tmpv[j] : tmpv[j];
            o += 3u;
          }
          for (uint32_t m = 0; m < nm; ++m) {
            double* xm = x + st * o;
            if (axial) { matmat_x_i(tmpm, xm, R, cplx); matmat_i_x(xm, iR, tmpm, cplx); }
            else { matmat_x_i(tmpm, xm, iR, cplx); matmat_i_x(xm, R, tmpm, cplx); }
            o += 9u;
          }
        }
        return BRILLE_OK;
      }
      return BRILLE_ERR_ROTATION;
    case BRILLE_LU_RECIPROCAL_LATTICE:
      if (d->rotates_like != BRILLE_ROT_VECTOR) return BRILLE_ERR_ROTATION;
      if (is_identity(R)) return BRILLE_OK;
      transpose_i(Rt, R);
      transpose_i(iRt, iR);
      for (uint32_t b = 0; b < B; ++b) {
        uint32_t o = b * s_ + ns;
        for (uint32_t v = 0; v < nv; ++v) {
          double* xv = x + st * o;
          if (cplx) mat_cvec_i(tmpv, Rt, xv); else mat_vec_real_i(tmpv, Rt, xv);
          for (int j = 0; j < 3 * st; ++j) xv[j] = tmpv[j];
          o += 3u;
        }
        for (uint32_t m = 0; m < nm; ++m) {
          double* xm = x + st * o;
          matmat_x_i(tmpm, xm, iRt, cplx);
          matmat_i_x(xm, Rt, tmpm, cplx);
          o += 9u;
        }
      }
      return BRILLE_OK;
    case BRILLE_LU_ANGSTROM:
      if (d->rotates_like != BRILLE_ROT_GAMMA) return BRILLE_ERR_ROTATION;
      break;
    default:
      return BRILLE_ERR_ROTATION;
  }
  /* Gamma: complex data required */
  if (!cplx || !gt) return BRILLE_ERR_ROTATION;
  const int cart = d->length_unit == BRILLE_LU_ANGSTROM;
  for (uint32_t b = 0; b < B; ++b) {
    uint32_t o = b * s_ + ns;
    if (nv > 0) {
      uint32_t o0 = o;
      for (uint32_t k = 0; k < nv; ++k) {
        double t0[6];
        if (cart) mat_cvec_d(t0, gt->cart_rotations + 9 * (uint64_t)ir, x + 2 * o);
        else mat_cvec_i(t0, iR, x + 2 * o);
        uint32_t key = k * gt->n_ops + ir;
        uint32_t v0idx = 3u * gt->F0[key];
        const double* dv = gt->vectors + 3 * (uint64_t)gt->vector_index[key];
        /* e_iqd  src/interpolator_gamma.tpp:18-32 */
        double dotqd = 0;
        for (int kk = 0; kk < 3; ++kk) dotqd += irq[kk] * dv[kk];
        double arg = TWO_PI * dotqd;
        double pr = cos(arg), pi = sin(arg);
        for (int j = 0; j < 3; ++j) {
          double ar = t0[2 * j], ai = t0[2 * j + 1];
          tA[2 * (v0idx + j)] = pr * ar - pi * ai;
          tA[2 * (v0idx + j) + 1] = pr * ai + pi * ar;
        }
        o += 3u;
      }
      for (uint32_t j = 0; j < nv * 3u; ++j) { x[2 * (o0 + j)] = tA[2 * j]; x[2 * (o0 + j) + 1] = tA[2 * j + 1]; }
    }
    /* Gamma transformation of 3N x 3N tensors is not exercised by this path */
  }
  return BRILLE_OK;
}

/* ------------------------------------------------------------------------ */
/* top level                                                                 */
/* ------------------------------------------------------------------------ */
/*
 * BrillouinZoneTrellis3::ir_interpolate_at  src/bz_trellis.hpp:153-196
 * values_out: N x B x span (double); vectors_out: N x B x span complex (interleaved).
 * Returns a brille_status; per-point status in aux->status when given.
 * no_move: 0 fold into the irreducible wedge; 1 do_not_move_points; 2 the
 * first-zone fold of BrillouinZoneTrellis3::interpolate_at (src/bz_trellis.hpp:
 * 111-128: moveinto, then the trellis interpolation with no rotation).
 */
/* wall seconds of the stages of the last brille_oracle_ir_interpolate_at on
 * this thread: (a) fold-in, (b)+(c1) location and sums, (c2) rotation (the
 * split of the reference's own profile, validation/profiling.md:85-93) */
static _Thread_local double g_stage_seconds[3];
static double now_seconds(void) { return omp_get_wtime(); }
void brille_oracle_stage_seconds(double* out) {
  for (int i = 0; i < 3; ++i) out[i] = g_stage_seconds[i];
}

int brille_oracle_ir_interpolate_at(const brille_grid_desc* g, const brille_data_desc* vals,
                                    const brille_data_desc* vecs, const brille_gamma_desc* gt,
                                    const brille_perm_desc* pd, const double* Q, uint64_t n,
                                    int no_move, int nthreads, double* values_out,
                                    double* vectors_out, const brille_aux_out* aux,
                                    uint64_t* first_bad) {
  if (!vals || !vecs || vals->n_vertices == 0) return BRILLE_ERR_NOT_FILLED;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  tol_t tbz = make_tols(g->float_tolerance, g->approx_digits);
  tol_t tdef = make_tols(0.0, 1);
  double* irq = (double*)malloc(sizeof(double) * 3 * (n ? n : 1));
  int32_t* tau = (int32_t*)malloc(sizeof(int32_t) * 3 * (n ? n : 1));
  uint32_t* rot = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* invrot = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint8_t* status = (uint8_t*)calloc(n ? n : 1, 1);
  int result = BRILLE_OK;
  uint64_t bad = n;

  double t_stage0 = now_seconds();
  /* stage (a): serial, src/bz_move.cpp:135-141, 262-285 */
  for (uint64_t i = 0; i < n; ++i) {
    if (no_move == 1) {
      for (int k = 0; k < 3; ++k) { irq[3 * i + k] = Q[3 * i + k]; tau[3 * i + k] = 0; }
      rot[i] = invrot[i] = g->identity_index;
      continue;
    }
    if (no_move == 2) {
      status[i] = (uint8_t)moveinto(g, Q + 3 * i, irq + 3 * i, tau + 3 * i, tbz);
      rot[i] = invrot[i] = g->identity_index;
      continue;
    }
    status[i] = (uint8_t)ir_moveinto(g, Q + 3 * i, irq + 3 * i, tau + 3 * i, rot + i, invrot + i, tbz, tdef);
  }

  const uint32_t B = vals->branches;
  const uint32_t sv = vals->elements[0] + vals->elements[1] + vals->elements[2];
  const uint32_t sc = vecs->elements[0] + vecs->elements[1] + vecs->elements[2];
  const uint64_t V = vals->n_vertices;
  memset(values_out, 0, sizeof(double) * n * B * sv);
  memset(vectors_out, 0, sizeof(double) * 2 * n * B * sc);

  double t_stage1 = now_seconds();
  /* stages (b)+(c1): src/trellis_poly.hpp:330 */
  int64_t sn = (int64_t)n;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
  for (int64_t si = 0; si < sn; ++si) {
    uint64_t i = (uint64_t)si;
    if (status[i] != BRILLE_PT_OK) continue;
    const double* q = irq + 3 * i;
    double x[3];
    mat_vec_d(x, g->outer_to_xyz, q); /* brille::get_xyz  src/array_lvec_methods.tpp:40-51 */
    uint64_t node = 0;
    if (!node_subscript(g, x, &node)) { status[i] = BRILLE_PT_NULL_NODE; continue; }
    uint32_t idx[8];
    double w[8];
    int nw = 0;
    uint32_t tet = 0xFFFFFFFFu;
    if (g->node_type[node] == BRILLE_NODE_CUBE)
      nw = cube_weights(g, g->node_index[node], x, idx, w, tdef);
    else
      nw = poly_weights(g, g->node_index[node], x, idx, w, &tet, tdef);
    if (aux && aux->node) aux->node[i] = (uint32_t)node;
    if (aux && aux->tet) aux->tet[i] = tet;
    if (aux && aux->n_iw) aux->n_iw[i] = (uint8_t)nw;
    for (int k = 0; k < 8; ++k) {
      if (aux && aux->iw_index) aux->iw_index[8 * i + k] = k < nw ? idx[k] : 0u;
      if (aux && aux->iw_weight) aux->iw_weight[8 * i + k] = k < nw ? w[k] : 0.0;
    }
    if (nw < 1) { status[i] = BRILLE_PT_NOT_FOUND; continue; }
    /* DualInterpolator::interpolate_at  src/interpolatordual.hpp:151-155, 374-382 */
    uint32_t pivot = idx[0];
    double* ov = values_out + i * B * sv;
    for (int xk = 0; xk < nw; ++xk) {
      const double* dx = vals->data + (uint64_t)idx[xk] * B * sv;
      for (uint32_t b = 0; b < B; ++b) {
        uint32_t p = perm_at(pd, V, B, pivot, idx[xk], b);
        for (uint32_t s = 0; s < sv; ++s) ov[b * sv + s] += w[xk] * dx[p * sv + s];
      }
    }
    double* oc = vectors_out + 2 * i * B * sc;
    const double* d0 = vecs->data + 2 * (uint64_t)idx[0] * B * sc;
    for (int xk = 0; xk < nw; ++xk) {
      const double* dx = vecs->data + 2 * (uint64_t)idx[xk] * B * sc;
      for (uint32_t b = 0; b < B; ++b) {
        uint32_t p = perm_at(pd, V, B, pivot, idx[xk], b);
        double er, ei;
        antiphase(sc, d0 + 2 * b * sc, dx + 2 * p * sc, &er, &ei);
        double wr = w[xk] * er, wi = w[xk] * ei; /* weights[x]*eith */
        for (uint32_t s = 0; s < sc; ++s) {
          double dr = dx[2 * (p * sc + s)], di = dx[2 * (p * sc + s) + 1];
          oc[2 * (b * sc + s)] += wr * dr - wi * di;
          oc[2 * (b * sc + s) + 1] += wr * di + wi * dr;
        }
      }
    }
  }

  double t_stage2 = now_seconds();
  /* stage (c2): src/bz_trellis.hpp:174-192 (interpolate_at has none) */
  int rot_status = BRILLE_OK;
#pragma omp parallel num_threads(nthreads) if (no_move != 2)
  if (no_move != 2) {
    uint32_t maxs = sc > sv ? sc : sv;
    double* tA = (double*)malloc(sizeof(double) * 2 * (maxs + 3));
    int local = BRILLE_OK;
#pragma omp for schedule(static)
    for (int64_t si = 0; si < sn; ++si) {
      uint64_t i = (uint64_t)si;
      if (status[i] != BRILLE_PT_OK) continue;
      int s1 = rotate_point(g, vals, gt, values_out + i * B * sv, irq + 3 * i, rot[i], invrot[i], tA);
      int s2 = rotate_point(g, vecs, gt, vectors_out + 2 * i * B * sc, irq + 3 * i, rot[i], invrot[i], tA);
      if (s1 != BRILLE_OK) local = s1;
      if (s2 != BRILLE_OK) local = s2;
    }
    free(tA);
    if (local != BRILLE_OK) {
#pragma omp critical
      rot_status = local;
    }
  }

  double t_stage3 = now_seconds();
  g_stage_seconds[0] = t_stage1 - t_stage0;
  g_stage_seconds[1] = t_stage2 - t_stage1;
  g_stage_seconds[2] = t_stage3 - t_stage2;
  /* The reference raises the first failure of the earliest stage: moveinto's
   * check over all points (src/bz_move.cpp:158), then the wedge search's first
   * point (:288-294), then a null node (rethrown, src/trellis_poly.hpp:341),
   * then the count of points not found (:345-349).  The point statuses are
   * numbered in that order: the lowest status wins, then the lowest index. */
  uint8_t worst = 255;
  for (uint64_t i = 0; i < n; ++i)
    if (status[i] != BRILLE_PT_OK && status[i] < worst) { worst = status[i]; bad = i; }
  if (worst != 255) {
    switch (worst) {
      case BRILLE_PT_OUTSIDE_BZ: result = BRILLE_ERR_OUTSIDE_BZ; break;
      case BRILLE_PT_OUTSIDE_IR: result = BRILLE_ERR_OUTSIDE_IR; break;
      case BRILLE_PT_NULL_NODE: result = BRILLE_ERR_NULL_NODE; break;
      default: result = BRILLE_ERR_NOT_FOUND;
    }
  }
  if (result == BRILLE_OK && rot_status != BRILLE_OK) result = rot_status;
  if (aux) {
    for (uint64_t i = 0; i < n; ++i) {
      if (aux->ir_q) for (int k = 0; k < 3; ++k) aux->ir_q[3 * i + k] = irq[3 * i + k];
      if (aux->tau) for (int k = 0; k < 3; ++k) aux->tau[3 * i + k] = tau[3 * i + k];
      if (aux->rot) aux->rot[i] = rot[i];
      if (aux->invrot) aux->invrot[i] = invrot[i];
      if (aux->status) aux->status[i] = status[i];
    }
  }
  if (first_bad) *first_bad = bad;
  free(irq); free(tau); free(rot); free(invrot); free(status);
  return result;
}

/* stage (a) alone, for tests of the fold-in: BrillouinZone::ir_moveinto */
int brille_oracle_ir_moveinto(const brille_grid_desc* g, const double* Q, uint64_t n,
                              double* q, int32_t* tau, uint32_t* rot, uint32_t* invrot,
                              uint8_t* status) {
  tol_t tbz = make_tols(g->float_tolerance, g->approx_digits);
  tol_t tdef = make_tols(0.0, 1);
  int result = BRILLE_OK;
  for (uint64_t i = 0; i < n; ++i) {
    rot[i] = invrot[i] = 0;
    int s = ir_moveinto(g, Q + 3 * i, q + 3 * i, tau + 3 * i, rot + i, invrot + i, tbz, tdef);
    status[i] = (uint8_t)s;
    if (s != BRILLE_PT_OK && result == BRILLE_OK)
      result = s == BRILLE_PT_OUTSIDE_BZ ? BRILLE_ERR_OUTSIDE_BZ : BRILLE_ERR_OUTSIDE_IR;
  }
  return result;
}

/* moveinto alone (first zone only): BrillouinZone::moveinto */
int brille_oracle_moveinto(const brille_grid_desc* g, const double* Q, uint64_t n, double* q,
                           int32_t* tau, uint8_t* status) {
  tol_t tbz = make_tols(g->float_tolerance, g->approx_digits);
  int result = BRILLE_OK;
  for (uint64_t i = 0; i < n; ++i) {
    int s = moveinto(g, Q + 3 * i, q + 3 * i, tau + 3 * i, tbz);
    status[i] = (uint8_t)s;
    if (s != BRILLE_PT_OK) result = BRILLE_ERR_OUTSIDE_BZ;
  }
  return result;
}

/* point location alone on xyz points (no fold-in): node_subscript + indices_weights */
int brille_oracle_locate(const brille_grid_desc* g, const double* xyz, uint64_t n, uint32_t* node,
                         uint32_t* tet, uint8_t* n_iw, uint32_t* iw_index, double* iw_weight) {
  tol_t tdef = make_tols(0.0, 1);
  int result = BRILLE_OK;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t nd = 0;
    uint32_t idx[8] = {0};
    double w[8] = {0};
    uint32_t t = 0xFFFFFFFFu;
    int nw = 0;
    if (node_subscript(g, xyz + 3 * i, &nd)) {
      if (g->node_type[nd] == BRILLE_NODE_CUBE) nw = cube_weights(g, g->node_index[nd], xyz + 3 * i, idx, w, tdef);
      else nw = poly_weights(g, g->node_index[nd], xyz + 3 * i, idx, w, &t, tdef);
    } else {
      result = BRILLE_ERR_NULL_NODE;
      nd = 0xFFFFFFFFu;
    }
    node[i] = (uint32_t)nd;
    tet[i] = t;
    n_iw[i] = (uint8_t)nw;
    for (int k = 0; k < 8; ++k) { iw_index[8 * i + k] = idx[k]; iw_weight[8 * i + k] = w[k]; }
  }
  return result;
}

/* the pieces the tests pin against closed forms and fixtures */
double brille_oracle_orient3d_lattice(const double* a, const double* b, const double* c,
                                      const double* d, const double* Grecip, const double* Greal,
                                      double scale) {
  return orient3d_lattice(a, b, c, d, Grecip, Greal, scale);
}
double brille_oracle_orient3d_xyz(const double* a, const double* b, const double* c, const double* d) {
  return orient3d_xyz(a, b, c, d);
}
int brille_oracle_approx(double a, double b, double Ttol, int tol) {
  return approx(a, b, make_tols(Ttol, tol));
}
uint32_t brille_oracle_find_bin(const double* edges, uint32_t n, double x) { return find_bin(edges, n, x); }
void brille_oracle_antiphase(uint32_t n, const double* a, const double* b, double* out) {
  antiphase(n, a, b, out, out + 1);
}

/* ------------------------------------------------------------------------ */
/* mode sorting: DualInterpolator::sort  src/interpolatordual.hpp:398-434    */
/* ------------------------------------------------------------------------ */
/* approx_float::scalar(act, 1.0) with its default tolerances, then the
 * clamp of |cos| to one (src/utilities.tpp:263-275, 369-373); returns 0 when
 * |cos| exceeds one by more than that (the reference throws there). */
static int clamp_cos(double* c) {
  double act = fabs(*c);
  if (approx(act, 1.0, make_tols(0.0, 1)) && act > 1) {
    *c /= act;
    act = fabs(*c);
  }
  return act <= 1;
}

/* utils::vector_angle, real  src/utilities.tpp:249-278 */
static double vec_angle_r(uint32_t n, const double* A, const double* B, int* ok) {
  double AA = 0, BB = 0, AB = 0;
  for (uint32_t i = 0; i < n; ++i) {
    AA += A[i] * A[i];
    BB += B[i] * B[i];
    AB += A[i] * B[i];
  }
  double nA = sqrt(AA), nB = sqrt(BB), c;
  if (nA != 0 && nB != 0) c = AB / (nA * nB);
  else c = (nA != 0 || nB != 0) ? 0.0 : 1.0;
  if (!clamp_cos(&c)) *ok = 0;
  return acos(c);
}

/* utils::euclidean_angle (vector_angle of complex vectors)  src/utilities.tpp:280-312 */
static double vec_angle_c(uint32_t n, const double* A, const double* B, int* ok) {
  double AB = 0, nA = 0, nB = 0, c;
  for (uint32_t i = 0; i < n; ++i) {
    AB += A[2 * i] * B[2 * i] + A[2 * i + 1] * B[2 * i + 1];
    nA += A[2 * i] * A[2 * i] + A[2 * i + 1] * A[2 * i + 1];
    nB += B[2 * i] * B[2 * i] + B[2 * i + 1] * B[2 * i + 1];
  }
  nA = sqrt(nA);
  nB = sqrt(nB);
  if (nA != 0 && nB != 0) c = AB / (nA * nB);
  else c = (nA != 0 || nB != 0) ? 0.0 : 1.0;
  if (!clamp_cos(&c)) *ok = 0;
  return acos(c);
}

/* utils::hermitian_product of complex vectors, sum of conj(a) b in element
 * order  src/utilities.tpp:340-350 */
static void herm_prod_c(uint32_t n, const double* a, const double* b, double* re, double* im) {
  double hr = 0, hi = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
    hr += ar * br + ai * bi;
    hi += ar * bi - ai * br;
  }
  *re = hr;
  *im = hi;
}

/* utils::vector_product, complex: |a^H b|^2  src/utilities.tpp:414-417 */
static double vec_prod_c(uint32_t n, const double* a, const double* b) {
  double hr, hi;
  herm_prod_c(n, a, b, &hr, &hi);
  return hr * hr + hi * hi;  /* real(h * conj(h)) */
}

/* utils::hermitian_angle, complex  src/utilities.tpp:357-381 */
static double herm_angle_c(uint32_t n, const double* A, const double* B, int* ok) {
  double nAB = sqrt(vec_prod_c(n, A, B)), hr, hi;
  herm_prod_c(n, A, A, &hr, &hi);
  double nA = sqrt(hr);
  herm_prod_c(n, B, B, &hr, &hi);
  double nB = sqrt(hr), c;
  if (nA != 0 && nB != 0) c = nAB / (nA * nB);
  else c = (nA != 0 || nB != 0) ? 0.0 : 1.0;
  if (!clamp_cos(&c)) *ok = 0;
  return acos(c);
}

/* utils::vector_distance  src/utilities.tpp:383-399 */
static double vec_dist_r(uint32_t n, const double* a, const double* b) {
  double s = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double d = a[i] - b[i];
    s += d * d;
  }
  return sqrt(s);
}
static double vec_dist_c(uint32_t n, const double* a, const double* b) {
  double s = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double dr = a[2 * i] - b[2 * i], di = a[2 * i + 1] - b[2 * i + 1];
    s += dr * dr + di * di;  /* real(d * conj(d)) */
  }
  return sqrt(s);
}

/* the vector cost functions  src/interpolator.hpp:256-293 */
static double vecfun_r(int t, uint32_t n, const double* a, const double* b, int* ok) {
  switch (t) {
    case 1: return vec_dist_r(n, a, b);
    case 2: {
      double h = 0;
      for (uint32_t i = 0; i < n; ++i) h += a[i] * b[i];
      return 1 - h;
    }
    case 3: return vec_angle_r(n, a, b, ok);
    case 4: return vec_angle_r(n, a, b, ok); /* hermitian_angle of real vectors */
    default: {
      double s = sin(vec_angle_r(n, a, b, ok));
      return s * s;
    }
  }
}
static double vecfun_c(int t, uint32_t n, const double* a, const double* b, int* ok) {
  switch (t) {
    case 1: return vec_dist_c(n, a, b);
    case 2: return 1 - vec_prod_c(n, a, b);
    case 3: return vec_angle_c(n, a, b, ok);
    case 4: return herm_angle_c(n, a, b, ok);
    default: {
      double s = sin(herm_angle_c(n, a, b, ok));
      return s * s;
    }
  }
}

/* Interpolator::add_cost for the real values, no phase freedom
 * src/interpolator_cost.tpp:42-57 */
static void add_cost_values(uint32_t B, const uint32_t* e, const double* x0, const double* x1,
                            const brille_cost_info* ci, double* cost, int* ok) {
  const uint32_t s_ = e[0] + e[1] + e[2], mo = e[0] + e[1];
  for (uint32_t i = 0; i < B; ++i) {
    const double* a = x0 + (size_t)i * s_;
    for (uint32_t j = 0; j < B; ++j) {
      const double* b = x1 + (size_t)j * s_;
      double s_cost = 0, v_cost = 0, m_cost = 0;
      if (e[0]) {
        s_cost = 0;
        for (uint32_t z = 0; z < e[0]; ++z) s_cost += fabs(a[z] - b[z]);
      }
      if (e[1]) v_cost = vecfun_r(ci->vector_cost, e[1], a + e[0], b + e[0], ok);
      if (e[2]) {
        m_cost = 0;
        for (uint32_t m = 0; m < e[2] / 9; ++m) m_cost += vec_dist_r(9, a + mo + 9 * m, b + mo + 9 * m);
      }
      cost[i * B + j] += ci->weights[0] * s_cost + ci->weights[1] * v_cost + ci->weights[2] * m_cost;
    }
  }
}

/* Interpolator::add_cost for the complex vectors, each pair compared after
 * the second is multiplied by its antiphase  src/interpolator_cost.tpp:26-41 */
static void add_cost_vectors(uint32_t B, const uint32_t* e, const double* x0, const double* x1,
                             const brille_cost_info* ci, double* cost, double* phased, int* ok) {
  const uint32_t s_ = e[0] + e[1] + e[2], mo = e[0] + e[1];
  for (uint32_t i = 0; i < B; ++i) {
    const double* a = x0 + 2 * (size_t)i * s_;
    for (uint32_t j = 0; j < B; ++j) {
      const double* b = x1 + 2 * (size_t)j * s_;
      double er, ei;
      antiphase(s_, a, b, &er, &ei);
      for (uint32_t z = 0; z < s_; ++z) { /* inplace_antiphase  src/utilities.tpp:589-593 */
        phased[2 * z] = er * b[2 * z] - ei * b[2 * z + 1];
        phased[2 * z + 1] = er * b[2 * z + 1] + ei * b[2 * z];
      }
      double s_cost = 0, v_cost = 0, m_cost = 0;
      if (e[0]) {
        s_cost = 0;
        for (uint32_t z = 0; z < e[0]; ++z) {
          double dr = a[2 * z] - phased[2 * z], di = a[2 * z + 1] - phased[2 * z + 1];
          s_cost += sqrt(dr * dr + di * di); /* utils::magnitude of a complex difference */
        }
      }
      if (e[1]) v_cost = vecfun_c(ci->vector_cost, e[1], a + 2 * e[0], phased + 2 * e[0], ok);
      if (e[2]) {
        m_cost = 0;
        for (uint32_t m = 0; m < e[2] / 9; ++m) m_cost += vec_dist_c(9, a + 2 * (mo + 9 * m), phased + 2 * (mo + 9 * m));
      }
      cost[i * B + j] += ci->weights[0] * s_cost + ci->weights[1] * v_cost + ci->weights[2] * m_cost;
    }
  }
}

/* Linear assignment: the n x n cost matrix C (row-major) is solved exactly
 * by successive shortest augmenting paths with row and column potentials
 * (Jonker and Volgenant, Computing 38, 325 (1987); the O(n^3) Hungarian
 * form of Kuhn and Munkres).  Rows are added in order 0..n-1; in each
 * Dijkstra step the unscanned column of least reduced distance is taken,
 * the lowest column index on ties.  row[i] = column of row i, col[j] = row
 * of column j.  For a cost matrix with a unique optimum every exact solver,
 * the reference's lapjv (src/lapjv.hpp) included, returns that optimum. */
static void lap_solve(uint32_t n, const double* C, int32_t* row, int32_t* col, double* u, double* v, double* minv,
                      int32_t* p, int32_t* way, uint8_t* used) {
  for (uint32_t j = 0; j <= n; ++j) {
    u[j] = 0;
    v[j] = 0;
    p[j] = 0;
    way[j] = 0;
  }
  for (uint32_t i = 1; i <= n; ++i) {
    p[0] = (int32_t)i;
    uint32_t j0 = 0;
    for (uint32_t j = 0; j <= n; ++j) {
      minv[j] = INFINITY;
      used[j] = 0;
    }
    do {
      used[j0] = 1;
      const uint32_t i0 = (uint32_t)p[j0];
      double delta = INFINITY;
      uint32_t j1 = 0;
      for (uint32_t j = 1; j <= n; ++j) {
        if (used[j]) continue;
        const double cur = C[(size_t)(i0 - 1) * n + (j - 1)] - u[i0] - v[j];
        if (cur < minv[j]) {
          minv[j] = cur;
          way[j] = (int32_t)j0;
        }
        if (minv[j] < delta) {
          delta = minv[j];
          j1 = j;
        }
      }
      for (uint32_t j = 0; j <= n; ++j) {
        if (used[j]) {
          u[p[j]] += delta;
          v[j] -= delta;
        } else {
          minv[j] -= delta;
        }
      }
      if (j1 == 0) return; /* no finite reduced cost left (NaN costs) */
      j0 = j1;
    } while (p[j0] != 0);
    do {
      const uint32_t j1 = (uint32_t)way[j0];
      p[j0] = p[j1];
      j0 = j1;
    } while (j0);
  }
  for (uint32_t j = 1; j <= n; ++j) {
    row[p[j] - 1] = (int32_t)(j - 1);
    col[j - 1] = p[j] - 1;
  }
}

int brille_oracle_lap(uint32_t n, const double* C, int32_t* row, int32_t* col) {
  double* buf = (double*)malloc(3 * (size_t)(n + 1) * sizeof(double));
  int32_t* ib = (int32_t*)malloc(2 * (size_t)(n + 1) * sizeof(int32_t));
  uint8_t* used = (uint8_t*)malloc(n + 1);
  if (!buf || !ib || !used) {
    free(buf);
    free(ib);
    free(used);
    return BRILLE_ERR_INVALID_ARGUMENT;
  }
  lap_solve(n, C, row, col, buf, buf + n + 1, buf + 2 * (n + 1), ib, ib + n + 1, used);
  free(buf);
  free(ib);
  free(used);
  return BRILLE_OK;
}

/* DualInterpolator::cost_matrix + determine_permutation_ij for the listed
 * pairs (i, j), i != j  src/interpolatordual.hpp:384-396, 422-432:
 * values->data is V x B x sv doubles, vectors->data V x B x sc complex.  For
 * pair k, row[k*B + a] is the mode of vertex j assigned to mode a of vertex
 * i (stored for key (i, j)), col[k*B + b] its inverse (key (j, i));
 * cost_out (nullable) receives the B x B matrices.  OpenMP over pairs as the
 * reference (src/interpolatordual.hpp:415-420). */
int brille_oracle_sort_pairs(uint64_t n_pairs, const uint32_t* pairs, const brille_data_desc* values,
                             const brille_cost_info* vcost, const brille_data_desc* vectors,
                             const brille_cost_info* ccost, int32_t* row, int32_t* col, double* cost_out,
                             int nthreads) {
  const uint32_t B = values->branches;
  if (vectors->branches != B || !vectors->is_complex || values->is_complex) return BRILLE_ERR_INVALID_ARGUMENT;
  const uint32_t sv = values->elements[0] + values->elements[1] + values->elements[2];
  const uint32_t sc = vectors->elements[0] + vectors->elements[1] + vectors->elements[2];
  int result = BRILLE_OK;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    double* cost = (double*)malloc((size_t)B * B * sizeof(double));
    double* phased = (double*)malloc(2 * (size_t)(sc ? sc : 1) * sizeof(double));
    double* buf = (double*)malloc(3 * (size_t)(B + 1) * sizeof(double));
    int32_t* ib = (int32_t*)malloc(2 * (size_t)(B + 1) * sizeof(int32_t));
    uint8_t* used = (uint8_t*)malloc(B + 1);
#pragma omp for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)n_pairs; ++k) {
      const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
      int ok = 1;
      for (uint32_t z = 0; z < B * B; ++z) cost[z] = 0;
      if (sv) add_cost_values(B, values->elements, values->data + (size_t)i * B * sv, values->data + (size_t)j * B * sv,
                              vcost, cost, &ok);
      if (sc) add_cost_vectors(B, vectors->elements, vectors->data + 2 * (size_t)i * B * sc,
                               vectors->data + 2 * (size_t)j * B * sc, ccost, cost, phased, &ok);
      if (!ok) {
#pragma omp critical
        result = BRILLE_ERR_NUMERIC;
      }
      if (cost_out) memcpy(cost_out + (size_t)k * B * B, cost, (size_t)B * B * sizeof(double));
      lap_solve(B, cost, row + (size_t)k * B, col + (size_t)k * B, buf, buf + B + 1, buf + 2 * (B + 1), ib,
                ib + B + 1, used);
    }
    free(cost);
    free(phased);
    free(buf);
    free(ib);
    free(used);
  }
  return result;
}
