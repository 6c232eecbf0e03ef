// Device-side tables and exact FP64 predicates of the ir_interpolate_at path.
//
// Every function here restates one reference routine with the same IEEE
// binary64 operation order; the library is compiled with -ffp-contract=off
// so no multiply-add is fused behind our back.  Citations are paths in the
// brille source tree.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

// The fold-in and location routines are host-callable too, so a CPU harness
// in tests/ can check the fast decisions against the exact ones.
#define BRILLE_HD __host__ __device__ inline

namespace brille_amd {

constexpr double kTwoPiD = 6.283185307179586476925286766559;
constexpr uint32_t kNoTet = 0xFFFFFFFFu;
constexpr int kMaxFaces = 32;

// Tables of one grid, resident in HBM.  Passed by value as a kernel argument.
struct GridConst {
  double outer_to_xyz[9];
  double outer_Gr[9];   // reciprocal metric
  double outer_Gd[9];   // real metric
  double outer_scale;   // det(B)/(2 pi)
  double prim_Gr[9];
  double prim_Gd[9];
  double prim_scale;
  int32_t P6t[9];
  int32_t invPt[9];
  int32_t transform_needed;
  int32_t no_ir_mirroring;
  int32_t has_seed;
  int32_t pad0;
  uint32_t F, W, G, identity;
  double tol_bz_rel, tol_bz_abs, tol_def_rel, tol_def_abs;
  // first zone, wedge, point group
  const double* prim_pa;
  const double* prim_pb;
  const double* prim_pc;
  const double* prim_n;
  const int32_t* prim_taus;
  const double* prim_tau_lens;
  const double* outer_pa;
  const double* outer_pb;
  const double* outer_pc;
  const double* wedge_n;
  const int32_t* rot;
  const uint32_t* inv_index;
  // fast-path tables (see fast_fold.hpp)
  const double* prim_aff;   // F x 4: affine form of the primitive face test
  const double* outer_aff;  // F x 4
  const double* wedge_aff;  // W x 3: metric-premultiplied wedge normals
  const double* orbit;      // G x 3: R_j p0 in the metric (score = q . orbit_j)
  double orbit_margin;
  double aff_margin;
  // trellis
  uint32_t nk[3];
  uint32_t pad1;
  const double* knots[3];
  double knot_lo[3];
  double knot_inv_step[3];
  const uint8_t* node_type;
  const uint32_t* node_index;
  const uint32_t* cube_v;
  const uint32_t* poly_off;
  const uint32_t* tet_v;
  const double* tet_c;
  const double* tet_vol;
  const double* verts;
  uint64_t n_nodes;
};

// Interpolation payload of one grid.
struct DataConst {
  uint64_t V;
  uint32_t B;
  uint32_t sv, sc;        // span of values / vectors per branch
  uint32_t ev[3], ec[3];  // element counts (scalars, vector elements, matrix elements)
  int32_t val_rot, val_lu, vec_rot, vec_lu;
  // One row per (vertex, branch): the sc vector elements (complex,
  // interleaved re/im) followed by the sv value elements, zero-padded to a
  // multiple of 16 bytes (to 32/64/128 bytes when the row fits in a cache
  // line), so a row never needs more than its own lines and loads 16 bytes at
  // a time.  Row (v, b) starts at rows + (v * B + b) * stride.
  const double* rows64;
  const float* rows32;
  uint32_t stride64, stride32;  // in reals
  // Row table (null until built; read by the bucketed sums only): for cube c,
  // slot x and mode b, row (c * 8 + x) * B + b, and for tetrahedron t row
  // tab_tet_row + (t * 4 + x) * B + b: the row of the slot-x vertex with every
  // vector element multiplied by its antiphase against the slot-0 vertex (the
  // pivot of every point whose slot-0 weight is non-zero), values as they
  // are; same stride as the payload rows.
  const float* cube_rows32;
  const double* cube_rows64;
  uint32_t tab_tet_row;
  // Wide-row payloads instead keep only the antiphases: the factor of slot x
  // for mode b of the block at table row trow (numbered as above) at
  // trow + x * B + b (null until built).
  const float2* wide_fac32;
  const double2* wide_fac64;
  // GammaTable
  uint32_t n_atoms, gamma_ops;
  const uint32_t* F0;
  const uint32_t* vidx;
  const double* gvec;
  // gvec[vidx[k * gamma_ops + op]] for every (atom k, op), 3 doubles each:
  // the phase vector of a key in one load instead of two dependent ones
  const double* gkey;
  const double* cart_rot;
  // PermutationTable
  uint64_t n_keys;
  const uint64_t* keys;
  const uint32_t* perm_index;
  const uint32_t* perms;
};

// Learned tetrahedron proposal: an MLP 3 -> 128 -> 128 -> K over the
// tetrahedra of the trellis (K = number of tetrahedra), evaluated only for the
// tetrahedra of the point's polyhedron node.  Inference weights are bf16
// (FP32 accumulation); W1 and the biases stay FP32.
constexpr int kPropH = 128;
struct ProposalConst {
  int32_t enabled;
  uint32_t K;
  const float* W1;     // H x 3
  const float* b1;     // H
  const uint16_t* W2;  // H x H bf16, row = output unit
  const float* b2;     // H
  const uint16_t* W3;  // K x H bf16, row c stored lane-major: element nb*16+m at m*8+nb
  const float* b3;     // K
  double margin;       // a proposal is taken when every barycentric weight exceeds it
  unsigned long long* stats;  // tried, verified
};

struct Tol {
  double rel, abs;
};

// approx_float::_scalar  src/approx_float.hpp:171-177
BRILLE_HD bool approx(double a, double b, Tol t) {
  double x = fabs(a - b);
  return x <= t.abs + t.rel * fabs(a + b) || x < DBL_MIN;
}
BRILLE_HD bool cmp_ge(double a, double b, Tol t) { return approx(a, b, t) || a > b; }
BRILLE_HD bool cmp_le(double a, double b, Tol t) { return approx(a, b, t) || a < b; }
BRILLE_HD bool cmp_gt(double a, double b, Tol t) { return !approx(a, b, t) && a > b; }

// utils::multiply_matrix_vector  src/utilities.tpp:44-55
BRILLE_HD void mat_vec(double* c, const double* A, const double* b) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += A[3 * i + k] * b[k];
    c[i] = s;
  }
}
BRILLE_HD void mat_vec_i(double* c, const int32_t* A, const double* b) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += (double)A[3 * i + k] * b[k];
    c[i] = s;
  }
}
// R^T b for an integer R (the wedge search's r_transpose product)
BRILLE_HD void mat_t_vec_i(double* c, const int32_t* R, const double* b) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += (double)R[3 * k + i] * b[k];
    c[i] = s;
  }
}

// same_lattice_dot  src/array_functions.hpp:238-246
BRILLE_HD double metric_dot(const double* x, const double* y, const double* M) {
  double t[3];
  mat_vec(t, M, x);
  double out = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) out += t[i] * y[i];
  return out;
}

// lattice pseudo_orient3d  src/geometry.hpp:118-138, src/array_functions.hpp:190-202,
// src/array_lvec_methods.tpp:53-66
BRILLE_HD double orient3d_lattice(const double* a, const double* b, const double* c,
                                                   const double* d, const double* Gr, const double* Gd,
                                                   double scale) {
  double u[3], v[3], w[3], cr[3], st[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) { u[i] = a[i] - d[i]; v[i] = b[i] - d[i]; w[i] = c[i] - d[i]; }
  cr[0] = v[1] * w[2] - v[2] * w[1];
  cr[1] = v[2] * w[0] - v[0] * w[2];
  cr[2] = v[0] * w[1] - v[1] * w[0];
#pragma unroll
  for (int i = 0; i < 3; ++i) cr[i] *= scale;
  mat_vec(st, Gd, cr);
#pragma unroll
  for (int i = 0; i < 3; ++i) st[i] /= kTwoPiD;
  return metric_dot(u, st, Gr);
}

// bare pseudo_orient3d on xyz  src/geometry.hpp:118-138
BRILLE_HD double orient3d_xyz(const double* a, const double* b, const double* c, const double* d) {
  double u[3], v[3], w[3], cr[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) { u[i] = a[i] - d[i]; v[i] = b[i] - d[i]; w[i] = c[i] - d[i]; }
  cr[0] = v[1] * w[2] - v[2] * w[1];
  cr[1] = v[2] * w[0] - v[0] * w[2];
  cr[2] = v[0] * w[1] - v[1] * w[0];
  double out = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) out += u[i] * cr[i];
  return out;
}

// point_inside_all_planes  src/geometry.hpp:413-419
BRILLE_HD bool inside_all_planes(uint32_t F, const double* pa, const double* pb, const double* pc,
                                                  const double* x, const double* Gr, const double* Gd, double scale,
                                                  Tol t) {
  double vmin = 0.0;
  for (uint32_t j = 0; j < F; ++j) {
    double a[3] = {pa[3 * j], pa[3 * j + 1], pa[3 * j + 2]};
    double b[3] = {pb[3 * j], pb[3 * j + 1], pb[3 * j + 2]};
    double c[3] = {pc[3 * j], pc[3 * j + 1], pc[3 * j + 2]};
    double o = orient3d_lattice(a, b, c, x, Gr, Gd, scale);
    if (j == 0 || o < vmin) vmin = o;
  }
  return vmin > 0 || approx(vmin, 0.0, t);
}

}  // namespace brille_amd
