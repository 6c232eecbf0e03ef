// Fused ir_interpolate_at tile kernel.
//
// One workgroup owns a tile of TILE points.
//   phase 1  (one lane per point): fold-in (stage a) and point location
//            (stage b) in FP64, in registers; the slots (vertex, weight), the
//            antiphase pivot and Ridx/invRidx go to LDS as a PointRec.
//   phase 2  (one lane per (point, mode) unit; consecutive lanes take
//            consecutive modes of a point, whose payload rows are adjacent):
//            weighted sums of values and of antiphased vectors (stage c1)
//            fused with the symmetry rotation (stage c2), written straight to
//            the output rows.
// Slots: a cube node fills 8 (slot i = vertex 7-i), a tetrahedron 4.  A slot
// the reference drops carries weight 0, so the non-zero slots in order are the
// reference's (index, weight) list and every loop over slots has a fixed trip
// count (no runtime-indexed arrays, hence no scratch memory).
// Real = double reproduces the reference's FP64 arithmetic; Real = float runs
// the sums and the rotation in FP32/complex64 with the same FP64 indices and
// weights.
#pragma once
#include <type_traits>

#include "exact_path.hpp"
#include "fast_fold.hpp"

namespace brille_amd {

constexpr int TILE = 256;

template <typename Real>
struct Cplx;
template <>
struct Cplx<double> {
  using T = double2;
  using V16 = double2;  // one 16-byte load
};
template <>
struct Cplx<float> {
  using T = float2;
  using V16 = float4;
};

// The first failure of a call, as the reference raises it: every point's
// fold-in runs before any location (src/bz_trellis.hpp:164-168), moveinto's
// check before the wedge search's (src/bz_move.cpp:158, 288-294), and a null
// node is rethrown before the count of points not found
// (src/trellis_poly.hpp:345-349).  The point statuses are numbered in that
// order, so the key (status << kErrStatusShift) | index, minimised over the
// failing points, gives the lowest status first and its lowest index.
constexpr int kErrStatusShift = 59;
struct ErrRec {
  unsigned long long first_bad;  // min over failing points of (status << kErrStatusShift) | index
  unsigned long long count;
};
__host__ __device__ __forceinline__ unsigned long long err_key(uint64_t index, int status) {
  return ((unsigned long long)status << kErrStatusShift) | (unsigned long long)index;
}

// Phase 1 result of one point, held in registers.
struct Located {
  double q[3];
  int32_t tau[3];
  uint32_t rot, invrot;
  uint32_t node, tet;
  uint32_t idx[8];
  double w[8];
  uint32_t pivot;  // first kept vertex: the antiphase reference of the sums
  uint32_t cube;   // cube index when the pivot is slot 0 (kNoCube otherwise)
  int nslot;       // 0 (failed), 4 or 8
  int nkept;
  int status;
};
constexpr uint32_t kNoCube = 0xFFFFFFFFu;

// What phase 2 needs of a point, in LDS.
template <typename Real>
struct PointRec {
  uint32_t idx[8];
  Real w[8];
  uint32_t pivot;
  uint32_t trow;  // row-table row of slot 0 for mode 0 (kNoCube: the point has no rows there)
  uint16_t rot, invrot;
  uint8_t nslot, pad0, pad1, pad2;
};

// Per-point constants of the Gamma rotation (rip_gamma_complex,
// src/interpolator_gamma.tpp:100-114), kept in LDS by phase 1 so that the sums
// read no tables: R is invRidx (integer, or Cartesian for angstrom units) as
// the sums' scale factors, to[a] the output atom of source atom a (GammaTable
// F0) and from[o] the source atom written to output atom o (255: none; the
// last source atom wins, as in a loop over source atoms).
template <typename Real>
struct GammaFrame {
  Real R[9];
  uint8_t to[4];
  uint8_t from[4];
};
template <typename Real>
constexpr int frame_words() {
  return (int)((sizeof(GammaFrame<Real>) + 7) / 8);
}

// ---------------------------------------------------------------------------
// payload rows

template <typename Real>
__device__ __forceinline__ const Real* payload_rows(const DataConst& d);
template <>
__device__ __forceinline__ const double* payload_rows<double>(const DataConst& d) { return d.rows64; }
template <>
__device__ __forceinline__ const float* payload_rows<float>(const DataConst& d) { return d.rows32; }
template <typename Real>
__device__ __forceinline__ uint32_t payload_stride(const DataConst& d) {
  return sizeof(Real) == 8 ? d.stride64 : d.stride32;
}
template <typename Real>
__device__ __forceinline__ const Real* row_at(const DataConst& d, uint32_t v, uint32_t b) {
  return payload_rows<Real>(d) + ((uint64_t)v * d.B + b) * payload_stride<Real>(d);
}

// the first N reals of a 16-byte aligned, 16-byte padded row
template <typename Real, int N>
__device__ __forceinline__ void load_row(Real* out, const Real* row) {
  using V = typename Cplx<Real>::V16;
  constexpr int PER = 16 / (int)sizeof(Real);
  constexpr int CH = (N + PER - 1) / PER;
  const V* p = reinterpret_cast<const V*>(row);
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    union {
      V v;
      Real r[PER];
    } u;
    u.v = p[c];
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (c * PER + j < N) out[c * PER + j] = u.r[j];
  }
}

// PermutationTable::safe_get with pivot iw[0]  src/permutation_table.hpp:193-198
__device__ __forceinline__ uint32_t perm_at(const DataConst& d, uint32_t pivot, uint32_t idx, uint32_t b) {
  if (d.n_keys == 0 || pivot == idx) return b;
  uint64_t key = (uint64_t)pivot * d.V + idx;
  uint64_t lo = 0, hi = d.n_keys;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (d.keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  if (lo < d.n_keys && d.keys[lo] == key) return d.perms[(uint64_t)d.perm_index[lo] * d.B + b];
  return b;
}

// ---------------------------------------------------------------------------
// antiphase

// one out-of-line copy of the reference's formula: the slot loops that call
// it are unrolled, and eight inlined atan2/sincos bodies would crowd the
// instruction cache
__device__ __noinline__ double2 antiphase_literal(double re, double im) {
  double th = -1.0 * atan2(im, re);
  double s, c;
  sincos(th, &s, &c);
  return make_double2(1.0 * c, 1.0 * s);
}

// utils::antiphase  src/utilities.tpp:567-577 of the Hermitian product
// z = re + i im: polar(1, -atan2(im, re)).  Literal keeps the reference's
// atan2/cos/sin; otherwise the identical closed form conj(z)/|z| (cos(atan2(y,
// x)) = x/|z|, sin(atan2(y,x)) = y/|z|), rescaled when |z|^2 would underflow or
// overflow, with the two cases atan2 fixes exactly kept exact: z real and
// positive gives 1, z = 0 gives +-1 by the sign of re.
template <typename Real, bool Literal>
__device__ __forceinline__ typename Cplx<Real>::T antiphase_of(Real re, Real im) {
  typename Cplx<Real>::T e;
  if constexpr (Literal) {
    const double2 t = antiphase_literal((double)re, (double)im);
    e.x = (Real)t.x;
    e.y = (Real)t.y;
  } else {
    constexpr Real lo = sizeof(Real) == 4 ? (Real)1e-36 : (Real)1e-300;
    constexpr Real hi = sizeof(Real) == 4 ? (Real)1e36 : (Real)1e300;
    const Real r2 = re * re + im * im;
    if (im == (Real)0 && re > (Real)0) {
      e.x = 1;
      e.y = 0;
    } else if (r2 > lo && r2 < hi) {
      Real inv;
      if constexpr (sizeof(Real) == 4) inv = __builtin_amdgcn_rsqf(r2);  // v_rsq_f32, 1 ulp
      else inv = 1.0 / sqrt(r2);
      e.x = re * inv;
      e.y = -im * inv;
    } else if (re != (Real)0 || im != (Real)0) {
      const Real m = fmax(fabs(re), fabs(im));
      const Real a = re / m, b = im / m;
      const Real inv = (Real)1 / sqrt(a * a + b * b);
      e.x = a * inv;
      e.y = -b * inv;
    } else {
      e.x = signbit(re) ? (Real)-1 : (Real)1;
      e.y = 0;
    }
  }
  return e;
}

// The closed form without branches, for the unrolled slot loops: re and im
// are rescaled by the power of two that brings the larger to [0.5, 1) (exact,
// and the result is the unscaled formula's wherever that one neither
// underflows nor overflows), so |z|^2 is in [0.25, 2] for the reciprocal
// square root; z real and positive gives exactly 1, z = 0 gives +-1 by the
// sign of re, as atan2 does.
template <typename Real>
__device__ __forceinline__ typename Cplx<Real>::T antiphase_closed(Real re, Real im) {
  typename Cplx<Real>::T e;
  const Real m = fmax(fabs(re), fabs(im));
  int k;
  (void)frexp(m, &k);
  const Real a = ldexp(re, -k), b = ldexp(im, -k);
  const Real r2 = a * a + b * b;
  Real inv;
  if constexpr (sizeof(Real) == 4) inv = __builtin_amdgcn_rsqf(r2);  // v_rsq_f32, 1 ulp
  else inv = 1.0 / sqrt(r2);
  Real ex = a * inv, ey = -b * inv;
  const bool zero = m == (Real)0;
  const bool pos_real = im == (Real)0 && re > (Real)0;
  ex = pos_real ? (Real)1 : (zero ? (signbit(re) ? (Real)-1 : (Real)1) : ex);
  ey = (pos_real || zero) ? (Real)0 : ey;
  e.x = ex;
  e.y = ey;
  return e;
}

// Hermitian product sum_k conj(p_k) x_k over n complex elements, in the
// reference's order (re += p.re x.re + p.im x.im; im += p.re x.im - p.im x.re)
template <typename Real>
__device__ __forceinline__ void hermitian(const Real* p, const Real* x, uint32_t n, Real* re, Real* im) {
  Real a = 0, b = 0;
  for (uint32_t k = 0; k < n; ++k) {
    a += p[2 * k] * x[2 * k] + p[2 * k + 1] * x[2 * k + 1];
    b += p[2 * k] * x[2 * k + 1] - p[2 * k + 1] * x[2 * k];
  }
  *re = a;
  *im = b;
}

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

__device__ __forceinline__ bool is_identity_rot(const int32_t* R) {
  return R[0] == 1 && R[1] == 0 && R[2] == 0 && R[3] == 0 && R[4] == 1 && R[5] == 0 && R[6] == 0 && R[7] == 0 &&
         R[8] == 1;
}

// element accessors so real and complex payloads share the rotation code
template <typename E>
struct Elem;
template <>
struct Elem<double> {
  __device__ static double zero() { return 0.0; }
  __device__ static double scale(double a, double s) { return a * s; }
  __device__ static double add(double a, double b) { return a + b; }
};
template <>
struct Elem<float> {
  __device__ static float zero() { return 0.f; }
  __device__ static float scale(float a, float s) { return a * s; }
  __device__ static float add(float a, float b) { return a + b; }
};
template <>
struct Elem<double2> {
  __device__ static double2 zero() { return make_double2(0.0, 0.0); }
  __device__ static double2 scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
  __device__ static double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
};
template <>
struct Elem<float2> {
  __device__ static float2 zero() { return make_float2(0.f, 0.f); }
  __device__ static float2 scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
  __device__ static float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
};

// c = A v, A integer or real 3x3 (mul_mat_vec  src/utilities.tpp:58-63)
template <typename E, typename M, typename S>
__device__ __forceinline__ void rot3(E* c, const M* A, const E* v) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    E s = Elem<E>::zero();
#pragma unroll
    for (int k = 0; k < 3; ++k) s = Elem<E>::add(s, Elem<E>::scale(v[k], (S)A[3 * i + k]));
    c[i] = s;
  }
}
template <typename E, typename S>
__device__ __forceinline__ void rot3_t(E* c, const int32_t* A, const E* v) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    E s = Elem<E>::zero();
#pragma unroll
    for (int k = 0; k < 3; ++k) s = Elem<E>::add(s, Elem<E>::scale(v[k], (S)A[3 * k + i]));
    c[i] = s;
  }
}
// C = X R and C = R X for a 3x3 element matrix X (mul_mat_mat  src/utilities.hpp:119)
template <typename E, typename S>
__device__ __forceinline__ void mm_xr(E* C, const E* X, const int32_t* R, bool Rt) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      E s = Elem<E>::zero();
#pragma unroll
      for (int k = 0; k < 3; ++k) s = Elem<E>::add(s, Elem<E>::scale(X[3 * i + k], (S)(Rt ? R[3 * j + k] : R[3 * k + j])));
      C[3 * i + j] = s;
    }
}
template <typename E, typename S>
__device__ __forceinline__ void mm_rx(E* C, const int32_t* R, const E* X, bool Rt) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      E s = Elem<E>::zero();
#pragma unroll
      for (int k = 0; k < 3; ++k) s = Elem<E>::add(s, Elem<E>::scale(X[3 * k + j], (S)(Rt ? R[3 * k + i] : R[3 * i + k])));
      C[3 * i + j] = s;
    }
}

// Rotation of one accumulated 3-vector / 3x3 matrix for the non-Gamma modes.
// rip_real src/interpolator_real.tpp:41-58, rip_recip src/interpolator_recip.tpp:41-55,
// rip_axial src/interpolator_axial.tpp:45-60.
template <typename E, typename S>
__device__ __forceinline__ void rotate_vec3(int rot_kind, int lu, const GridConst& g, uint32_t r, uint32_t ir, E* v) {
  const int32_t* R = g.rot + 9 * (uint64_t)r;
  if (is_identity_rot(R)) return;
  E t[3];
  if (lu == BRILLE_LU_RECIPROCAL_LATTICE) {
    rot3_t<E, S>(t, R, v);
  } else if (rot_kind == BRILLE_ROT_PSEUDOVECTOR) {
    const int32_t* iR = g.rot + 9 * (uint64_t)ir;
    rot3<E, int32_t, S>(t, iR, v);
    S det = (S)(R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) + R[2] * (R[3] * R[7] - R[4] * R[6]));
#pragma unroll
    for (int j = 0; j < 3; ++j) t[j] = Elem<E>::scale(t[j], det);
  } else {
    rot3<E, int32_t, S>(t, R, v);
  }
  v[0] = t[0]; v[1] = t[1]; v[2] = t[2];
}
template <typename E, typename S>
__device__ __forceinline__ void rotate_mat3(int rot_kind, int lu, const GridConst& g, uint32_t r, uint32_t ir, E* m) {
  const int32_t* R = g.rot + 9 * (uint64_t)r;
  if (is_identity_rot(R)) return;
  const int32_t* iR = g.rot + 9 * (uint64_t)ir;
  E t[9];
  if (lu == BRILLE_LU_RECIPROCAL_LATTICE) {
    mm_xr<E, S>(t, m, iR, true);
    mm_rx<E, S>(m, R, t, true);
  } else if (rot_kind == BRILLE_ROT_PSEUDOVECTOR) {
    mm_xr<E, S>(t, m, R, false);
    mm_rx<E, S>(m, iR, t, false);
  } else {
    mm_xr<E, S>(t, m, iR, false);
    mm_rx<E, S>(m, R, t, false);
  }
}

// e^{i 2 pi q . d} for atom k of a point  src/interpolator_gamma.tpp:18-32
template <typename Real>
__device__ inline typename Cplx<Real>::T gamma_phase(const DataConst& d, const double* q, uint32_t invrot, uint32_t k) {
  const uint32_t key = k * d.gamma_ops + invrot;
  const double* dv = d.gkey ? d.gkey + 3 * (uint64_t)key : d.gvec + 3 * (uint64_t)d.vidx[key];
  double dot = 0;
  dot += q[0] * dv[0];
  dot += q[1] * dv[1];
  dot += q[2] * dv[2];
  double arg = kTwoPiD * dot;
  typename Cplx<Real>::T ph;
  if constexpr (sizeof(Real) == 8) {
    double s, c;
    sincos(arg, &s, &c);
    ph.x = c;
    ph.y = s;
  } else {
    float s, c;
    sincosf((float)arg, &s, &c);
    ph.x = c;
    ph.y = s;
  }
  return ph;
}

// ---------------------------------------------------------------------------
// phase 2, general payload: any elements, rotations, GammaTable and
// PermutationTable; two passes over the rows (antiphases, then sums).

__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// complex element k of a row times its antiphase e (the values are not phased)
template <typename Real>
__device__ __forceinline__ void phase_elem(typename Cplx<Real>::T e, Real a, Real b, Real* px, Real* py) {
  *px = e.x * a - e.y * b;
  *py = e.x * b + e.y * a;
}

// One weighted term of a sum.  Literal (the plain kernel): acc + (w e) v, as
// the reference writes it; fast: the antiphased element e v rounded, then
// weighted with one fused multiply-add, which is unit_onepass's arithmetic.
template <typename Real, bool Literal>
__device__ __forceinline__ Real wsum_real(Real acc, Real w, Real v) {
  if constexpr (Literal) return acc + w * v;
  else return fma_r(w, v, acc);
}
// cw: w e (literal) or e (fast)
template <typename Real, bool Literal>
__device__ __forceinline__ typename Cplx<Real>::T wsum_cplx(typename Cplx<Real>::T acc, Real w,
                                                            typename Cplx<Real>::T cw, typename Cplx<Real>::T v) {
  if constexpr (Literal) {
    return Elem<typename Cplx<Real>::T>::add(acc, cmul(cw, v));
  } else {
    Real px, py;
    phase_elem<Real>(cw, v.x, v.y, &px, &py);
    acc.x = fma_r(w, px, acc.x);
    acc.y = fma_r(w, py, acc.y);
    return acc;
  }
}

template <typename Real, bool Literal>
__device__ inline void unit_generic(const GridConst& g, const DataConst& d, const PointRec<Real>& r, const double* q,
                                    uint32_t b, Real* ov, typename Cplx<Real>::T* oc,
                                    const typename Cplx<Real>::T* phases) {
  using C = typename Cplx<Real>::T;
  const uint32_t sv = d.sv, sc = d.sc;
  const uint32_t ns = d.ev[0], nv = d.ev[1] / 3u, nm = d.ev[2] / 9u;
  const int nslot = r.nslot;
  const Real* prow = row_at<Real>(d, r.pivot, b);
  const Real* row[8];
  Real w[8];
  C cw[8];
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const bool use = x < nslot && r.w[x] != (Real)0;
    w[x] = use ? r.w[x] : (Real)0;
    row[x] = use ? row_at<Real>(d, r.idx[x], perm_at(d, r.pivot, r.idx[x], b)) : prow;
    cw[x] = Elem<C>::zero();
    if (use && sc) {
      Real re, im;
      hermitian<Real>(prow, row[x], sc, &re, &im);
      C e = antiphase_of<Real, Literal>(re, im);
      if constexpr (Literal) {
        cw[x].x = w[x] * e.x;  // weights[x]*eith
        cw[x].y = w[x] * e.y;
      } else {
        cw[x] = e;  // the fast sums weight the antiphased element (unit_onepass's arithmetic)
      }
    }
  }
  // values  (interpolate_at_mix, real part)
  uint32_t e = 0;
  for (; e < ns; ++e) {
    Real acc = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x)
      if (w[x] != (Real)0) acc = wsum_real<Real, Literal>(acc, w[x], row[x][2 * sc + e]);
    ov[e] = acc;
  }
  for (uint32_t k = 0; k < nv; ++k, e += 3) {
    Real v[3] = {0, 0, 0};
#pragma unroll
    for (int x = 0; x < 8; ++x)
      if (w[x] != (Real)0)
#pragma unroll
        for (int j = 0; j < 3; ++j) v[j] = wsum_real<Real, Literal>(v[j], w[x], row[x][2 * sc + e + j]);
    rotate_vec3<Real, Real>(d.val_rot, d.val_lu, g, r.rot, r.invrot, v);
    ov[e] = v[0]; ov[e + 1] = v[1]; ov[e + 2] = v[2];
  }
  for (uint32_t k = 0; k < nm; ++k, e += 9) {
    Real m[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) m[j] = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x)
      if (w[x] != (Real)0)
#pragma unroll
        for (int j = 0; j < 9; ++j) m[j] = wsum_real<Real, Literal>(m[j], w[x], row[x][2 * sc + e + j]);
    rotate_mat3<Real, Real>(d.val_rot, d.val_lu, g, r.rot, r.invrot, m);
#pragma unroll
    for (int j = 0; j < 9; ++j) ov[e + j] = m[j];
  }
  // vectors (interpolate_at_mix, complex part)
  const uint32_t cs = d.ec[0], cv = d.ec[1] / 3u, cm = d.ec[2] / 9u;
  auto vec_at = [&](int x, uint32_t el) {
    C c;
    c.x = row[x][2 * el];
    c.y = row[x][2 * el + 1];
    return c;
  };
  e = 0;
  for (; e < cs; ++e) {
    C acc = Elem<C>::zero();
#pragma unroll
    for (int x = 0; x < 8; ++x)
      if (w[x] != (Real)0) acc = wsum_cplx<Real, Literal>(acc, w[x], cw[x], vec_at(x, e));
    oc[e] = acc;
  }
  const bool gamma = d.vec_rot == BRILLE_ROT_GAMMA;
  const bool cart = d.vec_lu == BRILLE_LU_ANGSTROM;
  for (uint32_t k = 0; k < cv; ++k, e += 3) {
    C v[3] = {Elem<C>::zero(), Elem<C>::zero(), Elem<C>::zero()};
#pragma unroll
    for (int x = 0; x < 8; ++x)
      if (w[x] != (Real)0)
#pragma unroll
        for (int j = 0; j < 3; ++j) v[j] = wsum_cplx<Real, Literal>(v[j], w[x], cw[x], vec_at(x, e + j));
    if (gamma) {
      // rip_gamma_complex  src/interpolator_gamma.tpp:100-114
      C t[3];
      if (cart) rot3<C, double, Real>(t, d.cart_rot + 9 * (uint64_t)r.invrot, v);
      else rot3<C, int32_t, Real>(t, g.rot + 9 * (uint64_t)r.invrot, v);
      const uint32_t l = d.F0[k * d.gamma_ops + r.invrot];
      C ph = phases ? phases[k] : gamma_phase<Real>(d, q, r.invrot, k);
      C* dst = oc + cs + 3u * l;
#pragma unroll
      for (int j = 0; j < 3; ++j) dst[j] = cmul(ph, t[j]);
    } else {
      rotate_vec3<C, Real>(d.vec_rot, d.vec_lu, g, r.rot, r.invrot, v);
      oc[e] = v[0]; oc[e + 1] = v[1]; oc[e + 2] = v[2];
    }
  }
  for (uint32_t k = 0; k < cm; ++k, e += 9) {
    C m[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) m[j] = Elem<C>::zero();
#pragma unroll
    for (int x = 0; x < 8; ++x)
      if (w[x] != (Real)0)
#pragma unroll
        for (int j = 0; j < 9; ++j) m[j] = wsum_cplx<Real, Literal>(m[j], w[x], cw[x], vec_at(x, e + j));
    rotate_mat3<C, Real>(d.vec_rot, d.vec_lu, g, r.rot, r.invrot, m);
#pragma unroll
    for (int j = 0; j < 9; ++j) oc[e + j] = m[j];
  }
}

// ---------------------------------------------------------------------------
// phase 2, specialised payloads with compile-time row sizes: SC complex vector
// elements and SV scalar values per (vertex, mode), no PermutationTable.
//   GAMMA: the vectors are SC/3 atom 3-vectors rotating like Gamma
//          (rip_gamma_complex  src/interpolator_gamma.tpp:100-114);
//   otherwise the vectors are scalars (no rotation).
// One pass: every row is read once, with 16-byte loads, and contributes its
// antiphase and its weighted elements before the next row is read.

template <typename Real>
__device__ __forceinline__ const Real* cube_rows(const DataConst& d);
template <>
__device__ __forceinline__ const float* cube_rows<float>(const DataConst& d) { return d.cube_rows32; }
template <>
__device__ __forceinline__ const double* cube_rows<double>(const DataConst& d) { return d.cube_rows64; }

// antiphased row of slot x for mode b of the block starting at table row trow
template <typename Real>
__device__ __forceinline__ const Real* table_row_at(const DataConst& d, uint32_t trow, int x, uint32_t b) {
  return cube_rows<Real>(d) + ((uint64_t)trow + (uint64_t)x * d.B + b) * payload_stride<Real>(d);
}

// The antiphase of slot x's row against the pivot row p: Hermitian product and
// closed form, one definition for the sums and for the cube row table.
template <typename Real, int SC>
__device__ __forceinline__ typename Cplx<Real>::T slot_antiphase(const Real* p, const Real* x) {
  Real re = 0, im = 0;
#pragma unroll
  for (int k = 0; k < SC; ++k) {
    re += p[2 * k] * x[2 * k] + p[2 * k + 1] * x[2 * k + 1];
    im += p[2 * k] * x[2 * k + 1] - p[2 * k + 1] * x[2 * k];
  }
  return antiphase_closed<Real>(re, im);
}

// Every one-pass sum is  acc += w * (e (x) row)  element by element, the
// product rounded first and the weight applied with one fused multiply-add:
// the cube row table stores the rounded products, so points summed from the
// table and points summed on the fly get the same bits.
// table: antiphased rows from the cube row table for the points it serves
template <typename Real, int SC, int SV, bool GAMMA, int GRP = 0>
__device__ __forceinline__ void unit_onepass(const DataConst& d, const PointRec<Real>& r, uint32_t b, Real* ov,
                                             typename Cplx<Real>::T* oc, const typename Cplx<Real>::T* phases,
                                             const GammaFrame<Real>* fr, bool table = false) {
  using C = typename Cplx<Real>::T;
  constexpr int NR = 2 * SC + SV;
  const Real* base = payload_rows<Real>(d);
  const uint64_t stride = payload_stride<Real>(d);
  // only cube points with the pivot in slot 0 have rows in the table
  const bool tab = table && r.trow != kNoCube;
  Real pr[2 * SC];
  if (!tab) load_row<Real, 2 * SC>(pr, base + ((uint64_t)r.pivot * d.B + b) * stride);
  Real acc[2 * SC];
  Real vacc[SV];
#pragma unroll
  for (int k = 0; k < 2 * SC; ++k) acc[k] = 0;
#pragma unroll
  for (int s = 0; s < SV; ++s) vacc[s] = 0;
  // The rows of a group of slots are all requested before any is used, so a
  // wave keeps GR rows per lane in flight (8 in FP32, 4 in FP64 for the
  // register budget).  Unused tetrahedron slots repeat the pivot with weight 0:
  // their loads hit in cache and add exact zeros, and no branch splits the loads.
  constexpr int GR = GRP ? GRP : (sizeof(Real) == 4 ? 8 : 4);
#pragma unroll
  for (int x0 = 0; x0 < 8; x0 += GR) {
    Real row[GR][NR];
    if (tab) {
#pragma unroll
      for (int x = 0; x < GR; ++x) {
        // a tetrahedron's block has four slots; the others weigh 0
        if (x0 + x < r.nslot) {
          load_row<Real, NR>(row[x], table_row_at<Real>(d, r.trow, x0 + x, b));
        } else {
#pragma unroll
          for (int k = 0; k < NR; ++k) row[x][k] = 0;
        }
      }
    } else {
#pragma unroll
      for (int x = 0; x < GR; ++x) load_row<Real, NR>(row[x], base + ((uint64_t)r.idx[x0 + x] * d.B + b) * stride);
#pragma unroll
      for (int x = 0; x < GR; ++x) {
        const C e = slot_antiphase<Real, SC>(pr, row[x]);
#pragma unroll
        for (int k = 0; k < SC; ++k)
          phase_elem<Real>(e, row[x][2 * k], row[x][2 * k + 1], &row[x][2 * k], &row[x][2 * k + 1]);
      }
    }
#pragma unroll
    for (int x = 0; x < GR; ++x) {
      const Real w = r.w[x0 + x];
#pragma unroll
      for (int k = 0; k < 2 * SC; ++k) acc[k] = fma_r(w, row[x][k], acc[k]);
#pragma unroll
      for (int s = 0; s < SV; ++s) vacc[s] = fma_r(w, row[x][2 * SC + s], vacc[s]);
    }
  }
#pragma unroll
  for (int s = 0; s < SV; ++s) ov[s] = vacc[s];
  if constexpr (GAMMA) {
#pragma unroll
    for (int a = 0; a < SC / 3; ++a) {
      C v[3], t[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        v[j].x = acc[2 * (3 * a + j)];
        v[j].y = acc[2 * (3 * a + j) + 1];
      }
      rot3<C, Real, Real>(t, fr->R, v);
      const uint32_t l = fr->to[a];
      const C ph = phases[a];
      C* dst = oc + 3u * l;
#pragma unroll
      for (int j = 0; j < 3; ++j) dst[j] = cmul(ph, t[j]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < SC; ++k) oc[k] = C{acc[2 * k], acc[2 * k + 1]};
  }
}

// ---------------------------------------------------------------------------
// Cooperative one-pass sums for the points the cube row table serves: G lanes
// per (point, mode) unit, lane q reading the 16-byte chunk q of each of the
// unit's eight antiphased rows.  A cube's rows for one slot lie mode after
// mode, so one wave load brings whole rows from a few cache lines instead of a
// 16-byte piece of each of 64 rows, and the sum is one fused multiply-add per
// element and slot.  Every output element gets unit_onepass's arithmetic in
// the same order, so both give the same bits.

// lanes per unit: the row stride in 16-byte chunks (1 = no cooperation)
template <typename Real, int SC, int SV>
constexpr int coop_lanes() {
  const int bytes = ((2 * SC + SV) * (int)sizeof(Real) + 15) / 16 * 16;
  if (bytes > 128) return 1;
  int p = 16;
  while (p < bytes) p <<= 1;
  return p / 16;
}

// the value of lane SRC of this lane's aligned group of G lanes
template <int G, int SRC>
__device__ __forceinline__ int grp_bcast_i(int v) {
  static_assert(G == 2 || G == 4 || G == 8, "groups of 2, 4 or 8 lanes");
  static_assert(SRC >= 0 && SRC < G, "source lane inside the group");
  if constexpr (G == 4) {
    return __builtin_amdgcn_mov_dpp(v, SRC * 0x55, 0xF, 0xF, false);
  } else if constexpr (G == 2) {
    return __builtin_amdgcn_mov_dpp(v, SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6), 0xF, 0xF, false);
  } else {
    // ds_swizzle bit mode: source lane = (lane & 0x18) | SRC within each half wave
    return __builtin_amdgcn_ds_swizzle(v, 0x18 | (SRC << 5));
  }
}
template <int G, int SRC>
__device__ __forceinline__ float grp_bcast(float v) {
  return __int_as_float(grp_bcast_i<G, SRC>(__float_as_int(v)));
}
template <int G, int SRC>
__device__ __forceinline__ double grp_bcast(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = grp_bcast_i<G, SRC>((int)(uint32_t)u);
  const int hi = grp_bcast_i<G, SRC>((int)(uint32_t)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}

// complex element M of the group's accumulator (lane M / CPL, pair M % CPL)
template <typename Real, int G, int CPL, int M>
__device__ __forceinline__ typename Cplx<Real>::T grp_elem(const Real* acc) {
  typename Cplx<Real>::T e;
  e.x = grp_bcast<G, (M / CPL) % G>(acc[2 * (M % CPL)]);
  e.y = grp_bcast<G, (M / CPL) % G>(acc[2 * (M % CPL) + 1]);
  return e;
}
template <typename Real, int G, int CPL, int SC, int M = 0>
__device__ __forceinline__ void grp_gather(const Real* acc, typename Cplx<Real>::T* all) {
  if constexpr (M < SC) {
    all[M] = grp_elem<Real, G, CPL, M>(acc);
    grp_gather<Real, G, CPL, SC, M + 1>(acc, all);
  }
}

// one lane's 16-byte chunk of a row
template <typename Real>
union Chunk {
  typename Cplx<Real>::V16 v;
  Real r[16 / sizeof(Real)];
};

// chunk q of the eight rows of a unit: off is the byte offset of chunk q of
// the slot-0 row in the table (below 4 GiB), slot_bytes the bytes from one
// slot's rows to the next (B rows)
template <typename Real>
__device__ __forceinline__ void coop_load(const DataConst& d, const PointRec<Real>& r, uint32_t off,
                                          uint32_t slot_bytes, Chunk<Real>* rows) {
  using V = typename Cplx<Real>::V16;
  const char* tb = reinterpret_cast<const char*>(cube_rows<Real>(d));
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    // a tetrahedron's block has four slots; the others weigh 0
    if (x < 4 || r.nslot == 8) {
      rows[x].v = *reinterpret_cast<const V*>(tb + (off + (uint32_t)x * slot_bytes));
    } else {
#pragma unroll
      for (int k = 0; k < (int)(16 / sizeof(Real)); ++k) rows[x].r[k] = 0;
    }
  }
}

// the weighted sum of a unit's chunks and the unit's outputs (see coop_load)
template <typename Real, int SC, int SV, bool GAMMA, int G>
__device__ __forceinline__ void unit_coop(const PointRec<Real>& r, const Chunk<Real>* rows, uint32_t q, Real* ov,
                                          typename Cplx<Real>::T* oc, const typename Cplx<Real>::T* phases,
                                          const GammaFrame<Real>* fr) {
  using C = typename Cplx<Real>::T;
  constexpr int CPL = 8 / (int)sizeof(Real);  // complex pairs per 16-byte chunk
  Real acc[2 * CPL];
#pragma unroll
  for (int k = 0; k < 2 * CPL; ++k) acc[k] = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const Real w = r.w[x];
#pragma unroll
    for (int k = 0; k < 2 * CPL; ++k) acc[k] = fma_r(w, rows[x].r[k], acc[k]);
  }
  // values: pair P >= SC holds values 2 (P - SC) and 2 (P - SC) + 1.  A
  // point's values are a few bytes written by the lanes of several groups:
  // plain stores, so that the L2 merges them (nontemporal 4-byte stores of
  // scattered points cost the sums 8 %).
#pragma unroll
  for (int h = 0; h < CPL; ++h) {
    const int vi = 2 * ((int)q * CPL + h - SC);
    if (vi >= 0 && vi < SV) ov[vi] = acc[2 * h];
    if (vi >= 0 && vi + 1 < SV) ov[vi + 1] = acc[2 * h + 1];
  }
  C out[CPL];
  bool put[CPL];
  if constexpr (GAMMA) {
    C all[SC];
    grp_gather<Real, G, CPL, SC>(acc, all);
#pragma unroll
    for (int h = 0; h < CPL; ++h) {
      const uint32_t m = q * CPL + h;
      out[h] = all[0];
      // output element m: component j of output atom m / 3, taken from the
      // source atom a the GammaTable maps there (none: unit_onepass leaves the
      // element unwritten too)
      const uint32_t oa = m / 3u, j = m - 3u * oa;
      const uint32_t a = m < (uint32_t)SC ? fr->from[oa < 4u ? oa : 0u] : 255u;
      put[h] = a != 255u;
      if (!put[h]) continue;
      C v[3];
      v[0] = v[1] = v[2] = all[0];
#pragma unroll
      for (int s = 0; s < SC / 3; ++s) {
        if ((uint32_t)s == a) {
          v[0] = all[3 * s];
          v[1] = all[3 * s + 1];
          v[2] = all[3 * s + 2];
        }
      }
      const Real* R = fr->R + 3 * j;
      C t = Elem<C>::zero();
#pragma unroll
      for (int k = 0; k < 3; ++k) t = Elem<C>::add(t, Elem<C>::scale(v[k], R[k]));
      out[h] = cmul(phases[a], t);
    }
  } else {
#pragma unroll
    for (int h = 0; h < CPL; ++h) {
      out[h] = C{acc[2 * h], acc[2 * h + 1]};
      put[h] = q * CPL + h < (uint32_t)SC;
    }
  }
  // the lane's elements in one 16-byte store where they are whole and aligned;
  // nontemporal, since a point's vectors are whole 16-byte pieces from
  // neighbouring lanes, not read again (plain stores here cost the sums 10 %)
  C* dst = oc + q * CPL;
  if constexpr (CPL == 2) {
    if (put[0] && put[1] && (reinterpret_cast<uintptr_t>(dst) & 15u) == 0) {
      typedef Real real4 __attribute__((ext_vector_type(4)));
      __builtin_nontemporal_store(real4{out[0].x, out[0].y, out[1].x, out[1].y}, reinterpret_cast<real4*>(dst));
      return;
    }
  }
  typedef Real real2 __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int h = 0; h < CPL; ++h)
    if (put[h]) __builtin_nontemporal_store(real2{out[h].x, out[h].y}, reinterpret_cast<real2*>(dst + h));
}

// Payload layouts with a specialised phase 2 (the general one handles all).
enum Layout : int {
  LAYOUT_GENERIC = 0,  // any elements, rotations and permutation table
  LAYOUT_GAMMA = 1,    // scalar values; vectors = one 3-vector per atom rotating like Gamma; no permutations
  LAYOUT_SCALAR = 2,   // scalar values and scalar vectors; no permutations
  LAYOUT_WIDE = 3      // as GAMMA or SCALAR with long rows (up to kWideMax complex): one wave per (point, mode)
};
constexpr uint32_t kWideMax = 128;  // complex elements per row the wide phase 2 handles (two per lane)
constexpr uint32_t kWideAtoms = kWideMax / 3;  // atoms of the longest wide Gamma row
// SC of the node-ordered wide sums (LAYOUT_WIDE): every point (kWideAll);
// only the atom-major walk, the points without a table block listed for a
// second launch (kWideAtom); that second launch over the listed points
// (kWideRest)
constexpr int kWideAll = 0, kWideAtom = 1, kWideRest = 2;
// LDS of the node-ordered wide sums' Gamma table: per-point phases of a chunk
// of the tile's points, every atom (FP32: 4608 complex, 192 points of 24
// atoms).  Those sums hold over 200 VGPRs, so two blocks per CU is their
// occupancy with or without it.
constexpr uint32_t kWidePhaseBytes = 36864;
// the same for the atom-only sums (kWideAtom below), whose 109 VGPRs allow
// four waves per SIMD when a block's LDS stays under 40 KB: 8 KB, 42 points
// of 24 atoms per chunk (with 36 KB, two waves: about 12 % slower)
constexpr uint32_t kWideAtomPhaseBytes = 8192;

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_sum_f(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// the wide factor table of a precision (antiphase of every slot row against
// its block's slot-0 row), or null
template <typename Real>
__device__ __forceinline__ const typename Cplx<Real>::T* wide_factors(const DataConst& d);
template <>
__device__ __forceinline__ const float2* wide_factors<float>(const DataConst& d) { return d.wide_fac32; }
template <>
__device__ __forceinline__ const double2* wide_factors<double>(const DataConst& d) { return d.wide_fac64; }

// Antiphases of NS slot rows against the pivot row for one wide-row unit,
// computed by the whole wave: lane l holds complex elements l and l + 64 of
// each row (p0/p1 of the pivot, v0/v1 of the slots); the Hermitian products
// are wave sums, all slots side by side.  One definition for the sums and
// for the factor table, so both give the same bits.
//
// FP32 reduces across the wave (a tree).  FP64 keeps the reference's order
// (src/utilities.tpp:567-577: one term per element, added to the running sum
// element by element): where two modes are degenerate their overlap is
// rounding noise around zero, and its antiphase is then decided by the exact
// order of the additions.  Each lane forms its elements' terms; the sums take
// them lane by lane with uniform readlane broadcasts.
__device__ __forceinline__ double readlane_f64(double v, uint32_t k) {
  const unsigned long long u = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, (int)k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), (int)k);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename Real, int NS>
__device__ __forceinline__ void wide_antiphases(typename Cplx<Real>::T p0, typename Cplx<Real>::T p1,
                                                const typename Cplx<Real>::T* v0, const typename Cplx<Real>::T* v1,
                                                typename Cplx<Real>::T* e, uint32_t sc) {
  if constexpr (sizeof(Real) == 8) {
    Real re0[NS], im0[NS], re1[NS], im1[NS], re[NS], im[NS];
#pragma unroll
    for (int x = 0; x < NS; ++x) {
      re0[x] = p0.x * v0[x].x + p0.y * v0[x].y;
      im0[x] = p0.x * v0[x].y - p0.y * v0[x].x;
      re1[x] = p1.x * v1[x].x + p1.y * v1[x].y;
      im1[x] = p1.x * v1[x].y - p1.y * v1[x].x;
      re[x] = 0;
      im[x] = 0;
    }
    const uint32_t n0 = sc < 64u ? sc : 64u;
    for (uint32_t k = 0; k < n0; ++k)
#pragma unroll
      for (int x = 0; x < NS; ++x) {
        re[x] += readlane_f64(re0[x], k);
        im[x] += readlane_f64(im0[x], k);
      }
    for (uint32_t k = 64; k < sc; ++k)
#pragma unroll
      for (int x = 0; x < NS; ++x) {
        re[x] += readlane_f64(re1[x], k - 64u);
        im[x] += readlane_f64(im1[x], k - 64u);
      }
#pragma unroll
    for (int x = 0; x < NS; ++x) e[x] = antiphase_closed<Real>(re[x], im[x]);
  } else {
    Real re[NS], im[NS];
#pragma unroll
    for (int x = 0; x < NS; ++x) {
      re[x] = p0.x * v0[x].x + p0.y * v0[x].y;
      re[x] += p1.x * v1[x].x + p1.y * v1[x].y;
      im[x] = p0.x * v0[x].y - p0.y * v0[x].x;
      im[x] += p1.x * v1[x].y - p1.y * v1[x].x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
      for (int x = 0; x < NS; ++x) {
        re[x] += __shfl_xor(re[x], o);
        im[x] += __shfl_xor(im[x], o);
      }
#pragma unroll
    for (int x = 0; x < NS; ++x) e[x] = antiphase_closed<Real>(re[x], im[x]);
  }
}

// What a lane's Γ rotation needs of a point, kept while a wave stays on it
template <typename Real>
struct WideCache {
  typename Cplx<Real>::T ph[2];
  Real R[2][3];
  uint32_t dst[2];
};

// The slot rows (lane l: complex elements l and l + 64 when TWO, value l) and
// antiphase factors (wide factor table when `tab`, else Hermitian products
// across the wave against the pivot row) of mode b of a point: what the sums
// of one wide unit read.  A point sharing the table block (trow) of the
// previous one has the same rows and factors.
template <typename Real, bool TWO>
struct WideRows {
  typename Cplx<Real>::T v0[8], v1[TWO ? 8 : 1], e[8];
  Real vv[8];
};

template <typename Real, bool TWO>
__device__ __forceinline__ void wide_load(const DataConst& d, const PointRec<Real>& r, uint32_t b, bool tab,
                                          WideRows<Real, TWO>& w) {
  using C = typename Cplx<Real>::T;
  const uint32_t l = threadIdx.x & 63u;
  const uint32_t sc = d.sc, sv = d.sv, B = d.B;
  const bool h0 = l < sc, h1 = TWO && l + 64u < sc;
  const C zero = Elem<C>::zero();
  const bool t = tab && r.trow != kNoCube;
  const C* fac = wide_factors<Real>(d);
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const Real* row = row_at<Real>(d, r.idx[x], b);
    const C* rc = reinterpret_cast<const C*>(row);
    w.v0[x] = h0 ? rc[l] : zero;
    if constexpr (TWO) w.v1[x] = h1 ? rc[l + 64] : zero;
    w.vv[x] = l < sv ? row[2 * sc + l] : (Real)0;
    if (t) w.e[x] = x < (int)r.nslot ? fac[r.trow + (uint32_t)x * B + b] : C{(Real)1, (Real)0};
  }
  if (!t) {
    const C* prow = reinterpret_cast<const C*>(row_at<Real>(d, r.pivot, b));
    const C p0 = h0 ? prow[l] : zero;
    C p1 = zero;
    if constexpr (TWO) p1 = h1 ? prow[l + 64] : zero;
    if constexpr (TWO) {
      wide_antiphases<Real, 8>(p0, p1, w.v0, w.v1, w.e, sc);
    } else {
      C z1[8];
#pragma unroll
      for (int x = 0; x < 8; ++x) z1[x] = zero;
      wide_antiphases<Real, 8>(p0, zero, w.v0, z1, w.e, sc);
    }
  }
}

// The rows of a WideRows multiplied by their antiphase factors in place
// (phase_elem, as the sums would per point): a table block's rows that
// several points reuse are antiphased once.
template <typename Real, bool TWO>
__device__ __forceinline__ void wide_phase_rows(WideRows<Real, TWO>& w) {
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    Real px, py;
    phase_elem<Real>(w.e[x], w.v0[x].x, w.v0[x].y, &px, &py);
    w.v0[x].x = px;
    w.v0[x].y = py;
    if constexpr (TWO) {
      phase_elem<Real>(w.e[x], w.v1[x].x, w.v1[x].y, &px, &py);
      w.v1[x].x = px;
      w.v1[x].y = py;
    }
  }
}

// The sums of mode b of a point from its loaded rows (PHASED: rows already
// antiphased by wide_phase_rows), the Gamma rotation through the wave's LDS
// slice `lds` (it mixes the three elements of an atom), and the stores.
// `wc` keeps the point's rotation data while `fresh` is false (the previous
// unit was the same point's).
template <typename Real, bool TWO, bool PHASED = false>
__device__ __forceinline__ void wide_sum_store(const GridConst& g, const DataConst& d, const PointRec<Real>& r,
                                               const double* q, uint32_t b, const WideRows<Real, TWO>& w,
                                               Real* ov_point, typename Cplx<Real>::T* oc_point,
                                               typename Cplx<Real>::T* lds, bool gamma, bool cart,
                                               WideCache<Real>& wc, bool fresh) {
  using C = typename Cplx<Real>::T;
  const uint32_t l = threadIdx.x & 63u;
  const uint32_t sc = d.sc, sv = d.sv;
  const bool h0 = l < sc, h1 = TWO && l + 64u < sc;
  const C zero = Elem<C>::zero();
  typedef Real real2 __attribute__((ext_vector_type(2)));
  C a0 = zero, a1 = zero;
  Real va = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const Real wx = r.w[x];
    Real px = w.v0[x].x, py = w.v0[x].y;
    if constexpr (!PHASED) phase_elem<Real>(w.e[x], w.v0[x].x, w.v0[x].y, &px, &py);
    a0.x = fma_r(wx, px, a0.x);
    a0.y = fma_r(wx, py, a0.y);
    if constexpr (TWO) {
      px = w.v1[x].x;
      py = w.v1[x].y;
      if constexpr (!PHASED) phase_elem<Real>(w.e[x], w.v1[x].x, w.v1[x].y, &px, &py);
      a1.x = fma_r(wx, px, a1.x);
      a1.y = fma_r(wx, py, a1.y);
    }
    va = fma_r(wx, w.vv[x], va);
  }
  Real* ov = ov_point + (uint64_t)b * sv;
  C* oc = oc_point + (uint64_t)b * sc;
  if (l < sv) __builtin_nontemporal_store(va, ov + l);
  if (!gamma) {
    if (h0) __builtin_nontemporal_store(real2{a0.x, a0.y}, reinterpret_cast<real2*>(oc + l));
    if (h1) __builtin_nontemporal_store(real2{a1.x, a1.y}, reinterpret_cast<real2*>(oc + l + 64));
    return;
  }
  // rip_gamma_complex  src/interpolator_gamma.tpp:100-114, element by element
  if (h0) lds[l] = a0;
  if (h1) lds[l + 64] = a1;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int h = 0; h < (TWO ? 2 : 1); ++h) {
    const uint32_t el = l + 64u * h;
    if (el < sc) {
      const uint32_t a = el / 3u, i = el - 3u * a;
      if (fresh) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
          wc.R[h][j] = cart ? (Real)d.cart_rot[9 * (uint64_t)r.invrot + 3 * i + j]
                            : (Real)g.rot[9 * (uint64_t)r.invrot + 3 * i + j];
        wc.ph[h] = gamma_phase<Real>(d, q, r.invrot, a);
        wc.dst[h] = 3u * d.F0[a * d.gamma_ops + r.invrot] + i;
      }
      const C v[3] = {lds[3 * a], lds[3 * a + 1], lds[3 * a + 2]};
      C t = zero;
#pragma unroll
      for (int j = 0; j < 3; ++j) t = Elem<C>::add(t, Elem<C>::scale(v[j], wc.R[h][j]));
      const C o = cmul(wc.ph[h], t);
      __builtin_nontemporal_store(real2{o.x, o.y}, reinterpret_cast<real2*>(oc + wc.dst[h]));
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The Gamma data of the node-ordered wide sums in LDS: phases e^{i 2 pi q.d}
// of a chunk of the tile's points (point-major, `cap` complex), and per
// point-group operation the rotation rows (9 per operation) and the output
// atom of every source atom.  All null: the sums compute them per unit.
template <typename Real>
struct WideGammaLds {
  typename Cplx<Real>::T* ph = nullptr;
  Real* rot = nullptr;
  uint8_t* dst = nullptr;
  uint32_t cap = 0;
};

// Fill the per-operation part of WideGammaLds (operations invrot < G); the
// caller synchronises the block before the sums read it.
template <typename Real, int TB>
__device__ __forceinline__ void wide_gamma_ops(const GridConst& g, const DataConst& d, const WideGammaLds<Real>& t) {
  const bool cart = d.vec_lu == BRILLE_LU_ANGSTROM;
  const uint32_t na = d.sc / 3u;
  for (uint32_t k = threadIdx.x; k < 9u * g.G; k += TB)
    t.rot[k] = cart ? (Real)d.cart_rot[k] : (Real)g.rot[k];
  for (uint32_t k = threadIdx.x; k < na * g.G; k += TB) {
    const uint32_t op = k / na, a = k - op * na;
    t.dst[k] = (uint8_t)d.F0[a * d.gamma_ops + op];
  }
}

// Phase 2 for rows too long for one lane: mode b of one point by one wave.
// Lane l takes the row's complex elements l and l + 64 (TWO: rows of more
// than 64), so every row is read with coalesced loads.
template <typename Real, bool TWO>
__device__ inline void unit_wide(const GridConst& g, const DataConst& d, const PointRec<Real>& r, const double* q,
                                 uint32_t b, Real* ov_point, typename Cplx<Real>::T* oc_point,
                                 typename Cplx<Real>::T* lds, bool gamma, bool cart, WideCache<Real>& wc, bool fresh,
                                 bool table) {
  WideRows<Real, TWO> w;
  wide_load<Real, TWO>(d, r, b, table, w);
  wide_sum_store<Real, TWO>(g, d, r, q, b, w, ov_point, oc_point, lds, gamma, cart, wc, fresh);
}

struct LaunchArgs {
  const double* Q;
  uint64_t n;
  uint64_t base;  // index of Q[0] in the caller's full array (for error reporting)
  void* vals;
  void* vecs;
  brille_aux_out aux;
  ErrRec* err;
  int fold;  // Fold
  int use_proposal;
  ProposalConst prop;
};

// ---------------------------------------------------------------------------
// phase 1 for one point

// How stage (a) folds a point: into the irreducible wedge (ir_interpolate_at),
// not at all (do_not_move_points), or into the first zone only, with the
// identity operation (interpolate_at, src/bz_trellis.hpp:111-128: the identity
// rotation and a zero Gamma phase leave the sums exactly as they are).
enum Fold : int { FOLD_IR = 0, FOLD_NONE = 1, FOLD_FIRST_ZONE = 2 };

// stage (a) and the node subscript of stage (b): q, tau, Ridx/invRidx, node
template <bool Fast>
__device__ inline void fold_node(const GridConst& g, const double* Q, int fold, Located& L) {
  const Tol tbz{g.tol_bz_rel, g.tol_bz_abs};
  const Tol tdef{g.tol_def_rel, g.tol_def_abs};
  L.tau[0] = L.tau[1] = L.tau[2] = 0;
  L.rot = L.invrot = g.identity;
  L.node = 0xFFFFFFFFu;
  L.tet = kNoTet;
  L.nslot = 0;
  L.nkept = 0;
  L.pivot = 0;
  L.status = BRILLE_PT_OK;
  if (fold == FOLD_NONE) {
    L.q[0] = Q[0]; L.q[1] = Q[1]; L.q[2] = Q[2];
  } else {
    int st = Fast ? fast_moveinto(g, Q, L.q, L.tau, tbz) : moveinto_exact(g, Q, L.q, L.tau, tbz);
    if (st == BRILLE_PT_OK && fold == FOLD_IR)
      st = Fast ? fast_wedge_search(g, L.q, &L.rot, &L.invrot, tbz, tdef)
                : wedge_search_exact(g, L.q, &L.rot, &L.invrot, tbz, tdef);
    if (st != BRILLE_PT_OK) { L.status = st; return; }
  }
  double x[3];
  mat_vec(x, g.outer_to_xyz, L.q);  // brille::get_xyz
  uint64_t node = 0;
  if (!node_subscript<Fast>(g, x, &node, tdef)) { L.status = BRILLE_PT_NULL_NODE; return; }
  L.node = (uint32_t)node;
}

// the rest of stage (b) for a folded point with a usable node: slots, pivot
// hint: a proposed tetrahedron of the node (kNoTet: none).  It is taken only
// when the reference's scan would stop there too: the circumsphere shortcut
// lets it through and every barycentric weight exceeds `margin` (a point that
// deep inside one tetrahedron is outside every other by more than the
// comparison tolerance, so no earlier tetrahedron passes the containment test).
// Otherwise the exact scan runs.  *verified tells which happened.
__device__ inline void node_slots(const GridConst& g, Located& L, uint32_t hint = kNoTet, double margin = 0.0,
                                  bool* verified = nullptr) {
  const Tol tdef{g.tol_def_rel, g.tol_def_abs};
  double x[3];
  mat_vec(x, g.outer_to_xyz, L.q);  // brille::get_xyz
  const uint32_t node = L.node;
  L.tet = kNoTet;
  L.cube = kNoCube;
  L.nslot = 0;
  L.nkept = 0;
  L.pivot = 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    L.idx[s] = 0;
    L.w[s] = 0.0;
  }
  if (g.node_type[node] == BRILLE_NODE_CUBE) {
    cube_slots(g, g.node_index[node], x, L.idx, L.w, tdef);
    L.nslot = 8;
    if (L.w[0] != 0.0) L.cube = g.node_index[node];
  } else {
    bool ok = false;
    if (hint != kNoTet && tet_might_contain(g, hint, x, tdef) >= 0.0) {
      double ww[4];
      tet_contains(g, hint, x, ww, false, tdef);
      ok = ww[0] > margin && ww[1] > margin && ww[2] > margin && ww[3] > margin;
      if (ok) {
        tet_slots(g, hint, ww, L.idx, L.w, tdef);
        L.tet = hint;
      }
    }
    if (verified) *verified = ok;
    if (!ok) L.tet = poly_slots(g, g.node_index[node], x, L.idx, L.w, tdef);
    L.nslot = L.tet == kNoTet ? 0 : 4;
  }
  // pivot = first kept slot; unused tetrahedron slots repeat it with weight 0
#pragma unroll
  for (int s = 7; s >= 0; --s)
    if (s < L.nslot && L.w[s] != 0.0) L.pivot = L.idx[s];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if (s < L.nslot && L.w[s] != 0.0) ++L.nkept;
    if (s >= L.nslot) L.idx[s] = L.pivot;
  }
  if (L.nkept < 1) { L.status = BRILLE_PT_NOT_FOUND; L.nslot = 0; }
}

template <bool Fast>
__device__ inline void locate(const GridConst& g, const double* Q, int fold, Located& L) {
  fold_node<Fast>(g, Q, fold, L);
  if (L.status == BRILLE_PT_OK) node_slots(g, L);
}

// aux outputs of one point (weights in FP64 whatever the sums' precision)
__device__ inline void write_aux(const brille_aux_out& aux, uint64_t p, const Located& L) {
  if (aux.ir_q) { aux.ir_q[3 * p] = L.q[0]; aux.ir_q[3 * p + 1] = L.q[1]; aux.ir_q[3 * p + 2] = L.q[2]; }
  if (aux.tau) { aux.tau[3 * p] = L.tau[0]; aux.tau[3 * p + 1] = L.tau[1]; aux.tau[3 * p + 2] = L.tau[2]; }
  if (aux.rot) aux.rot[p] = L.rot;
  if (aux.invrot) aux.invrot[p] = L.invrot;
  if (aux.node) aux.node[p] = L.node;
  if (aux.tet) aux.tet[p] = L.tet;
  if (aux.n_iw) aux.n_iw[p] = (uint8_t)(L.status == BRILLE_PT_OK ? L.nkept : 0);
  if (aux.iw_index || aux.iw_weight) {
    uint32_t k = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s)
      if (L.status == BRILLE_PT_OK && s < L.nslot && L.w[s] != 0.0) {
        if (aux.iw_index) aux.iw_index[8 * p + k] = L.idx[s];
        if (aux.iw_weight) aux.iw_weight[8 * p + k] = L.w[s];
        ++k;
      }
    for (; k < 8; ++k) {
      if (aux.iw_index) aux.iw_index[8 * p + k] = 0u;
      if (aux.iw_weight) aux.iw_weight[8 * p + k] = 0.0;
    }
  }
  if (aux.status) aux.status[p] = (uint8_t)L.status;
}

// ---------------------------------------------------------------------------
// shared pieces of the tile kernels.  L selects phase 2; SC/SV are the
// compile-time row sizes of the specialised layouts (ignored by
// LAYOUT_GENERIC).

template <int L, int SC, int TB = TILE, typename Real = double>
struct TileShape {
  static constexpr int NPH = L == LAYOUT_GAMMA ? SC / 3 : (L == LAYOUT_GENERIC ? 4 : 1);  // phases per point
  // per-point side data in s_q: q (general and wide layouts) or the GammaFrame
  static constexpr int NQ = (L == LAYOUT_GENERIC || L == LAYOUT_WIDE) ? 3 * TB
                            : (L == LAYOUT_GAMMA ? frame_words<Real>() * TB : 1);
  static constexpr int NW = L == LAYOUT_WIDE ? (TB / 64) * (int)kWideMax : 1;              // wide-row slices
};

// The row-table block of a point: cube c's eight slots at rows c * 8 * B, the
// four slots of tetrahedron t at tab_tet_row + t * 4 * B; only points whose
// antiphase pivot is slot 0 have one.
__device__ __forceinline__ uint32_t table_row(const DataConst& d, const Located& P) {
  if (P.cube != kNoCube) return P.cube * 8u * d.B;
  if (P.tet != kNoTet && P.nslot == 4 && P.w[0] != 0.0) return d.tab_tet_row + P.tet * 4u * d.B;
  return kNoCube;
}

// the Gamma frame of point-group operation invrot (it depends on nothing else)
template <typename Real, int SC>
__device__ __forceinline__ void fill_frame(const GridConst& g, const DataConst& d, uint32_t invrot,
                                           GammaFrame<Real>& f) {
  if (d.vec_lu == BRILLE_LU_ANGSTROM) {
#pragma unroll
    for (int k = 0; k < 9; ++k) f.R[k] = (Real)d.cart_rot[9 * (uint64_t)invrot + k];
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) f.R[k] = (Real)g.rot[9 * (uint64_t)invrot + k];
  }
  uint8_t to[4] = {255, 255, 255, 255}, from[4] = {255, 255, 255, 255};
#pragma unroll
  for (int a = 0; a < SC / 3; ++a) {
    const uint32_t l = d.F0[a * d.gamma_ops + invrot];
    to[a] = (uint8_t)l;
#pragma unroll
    for (int o = 0; o < SC / 3; ++o)
      if ((uint32_t)o == l) from[o] = (uint8_t)a;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    f.to[a] = to[a];
    f.from[a] = from[a];
  }
}

// a located point's phase-2 record (plus phases, and q for the general
// layout; frame = false: the Gamma frame comes from a per-operation table)
template <typename Real, int L, int SC>
__device__ __forceinline__ void store_record(const GridConst& g, const DataConst& d, const Located& P,
                                             PointRec<Real>& r, typename Cplx<Real>::T* s_ph, double* s_q,
                                             uint32_t slot, bool ph_in_lds, uint32_t n_atoms, bool frame = true) {
  constexpr int NPH = TileShape<L, SC>::NPH;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    r.idx[s] = P.idx[s];
    r.w[s] = (Real)P.w[s];
  }
  r.pivot = P.pivot;
  r.trow = table_row(d, P);
  r.rot = (uint16_t)P.rot;
  r.invrot = (uint16_t)P.invrot;
  r.nslot = (uint8_t)P.nslot;
  if constexpr (L == LAYOUT_GAMMA) {
#pragma unroll
    for (int k = 0; k < NPH; ++k) s_ph[NPH * slot + k] = gamma_phase<Real>(d, P.q, P.invrot, k);
    if (frame) fill_frame<Real, SC>(g, d, P.invrot, reinterpret_cast<GammaFrame<Real>*>(s_q)[slot]);
  } else if constexpr (L == LAYOUT_GENERIC) {
    s_q[3 * slot] = P.q[0]; s_q[3 * slot + 1] = P.q[1]; s_q[3 * slot + 2] = P.q[2];
    if (ph_in_lds)
      for (uint32_t k = 0; k < n_atoms; ++k) s_ph[NPH * slot + k] = gamma_phase<Real>(d, P.q, P.invrot, k);
  } else if constexpr (L == LAYOUT_WIDE) {
    s_q[3 * slot] = P.q[0]; s_q[3 * slot + 1] = P.q[1]; s_q[3 * slot + 2] = P.q[2];
  }
}

// phase 2 over the npts records of a tile of TB points (TB threads): point pl
// writes output row out_row[pl] (or row0 + pl when out_row is null); GRP
// overrides the rows the one-pass sums keep in flight (0: their default)
// table: the cube factor table is built; its points are the first n_coop
// records, summed by unit_coop where the rows allow it
template <typename Real, bool Fast, int L, int SC, int SV, int TB = TILE, int GRP = 0>
__device__ __forceinline__ void tile_sums(const GridConst& g, const DataConst& d, const PointRec<Real>* s_rec,
                                          const typename Cplx<Real>::T* s_ph, const double* s_q,
                                          const uint32_t* out_row, uint64_t row0, uint32_t npts, bool ph_in_lds,
                                          void* vals_, void* vecs_, typename Cplx<Real>::T* s_wide = nullptr,
                                          bool table = false, uint32_t n_coop = 0,
                                          const GammaFrame<Real>* op_frames = nullptr, bool node_order = false,
                                          WideGammaLds<Real> wg = {}, uint32_t b_lo = 0, uint32_t b_hi = ~0u) {
  using C = typename Cplx<Real>::T;
  constexpr int NPH = TileShape<L, SC>::NPH;
  const uint32_t B = d.B;
  const uint32_t units = npts * B;
  Real* vals = (Real*)vals_;
  C* vecs = (C*)vecs_;
  const bool cart = d.vec_lu == BRILLE_LU_ANGSTROM;
  if constexpr (L == LAYOUT_WIDE) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const bool gamma = d.vec_rot == BRILLE_ROT_GAMMA && d.ec[1] > 0;
    C* lds = s_wide + wave * kWideMax;
    WideCache<Real> wc;
    const bool two = d.sc > 64;
    const bool tab = wide_factors<Real>(d) != nullptr;
    if (!node_order && gamma && tab && d.sv <= d.sc / 3u) {
      // Point-major, atom-major (Gamma rows with the factor table): wave w
      // takes units w, w + TB/64, ... of (point, pass), a pass being 64 /
      // n_atoms modes with a lane per (mode, atom) as in the node-ordered
      // walk below; the wave keeps the point's phase, rotation rows and
      // output atom while it stays on the point.  Points without a table
      // block take the wave-per-unit sums for the pass's modes.
      const uint32_t na = d.sc / 3u, sc = d.sc, sv = d.sv;
      const uint32_t mpw = 64u / na;
      const uint32_t m = lane / na, a = lane - m * na;
      const uint32_t passes = (B + mpw - 1) / mpw;
      const C* fac = wide_factors<Real>(d);
      uint32_t last = 0xFFFFFFFFu;
      C ph = Elem<C>::zero();
      Real R[9];
      uint32_t dsta = 0;
      for (uint32_t c = wave; c < npts * passes; c += TB / 64) {
        const uint32_t pl = c / passes, t = c - pl * passes;
        const PointRec<Real>& r = s_rec[pl];
        const uint64_t row = ((out_row ? (uint64_t)out_row[pl] : row0 + pl)) * B;
        if (r.nslot == 0 || r.trow == kNoCube) {
          Real* ovp = vals + row * sv;
          C* ocp = vecs + row * sc;
          const uint32_t bend = (t + 1) * mpw < B ? (t + 1) * mpw : B;
          for (uint32_t b = t * mpw; b < bend; ++b) {
            if (r.nslot == 0) {
              for (uint32_t e = lane; e < sv; e += 64) ovp[(uint64_t)b * sv + e] = 0;
              for (uint32_t e = lane; e < sc; e += 64) ocp[(uint64_t)b * sc + e] = Elem<C>::zero();
            } else if (two) {
              unit_wide<Real, true>(g, d, r, s_q + 3 * pl, b, ovp, ocp, lds, gamma, cart, wc, true, tab);
            } else {
              unit_wide<Real, false>(g, d, r, s_q + 3 * pl, b, ovp, ocp, lds, gamma, cart, wc, true, tab);
            }
          }
          continue;
        }
        if (pl != last) {
          last = pl;
          if (m < mpw) {
            ph = gamma_phase<Real>(d, s_q + 3 * pl, r.invrot, a);
#pragma unroll
            for (int k = 0; k < 9; ++k)
              R[k] = cart ? (Real)d.cart_rot[9 * (uint64_t)r.invrot + k] : (Real)g.rot[9 * (uint64_t)r.invrot + k];
            dsta = d.F0[a * d.gamma_ops + r.invrot];
          }
        }
        const uint32_t b = t * mpw + m;
        if (m >= mpw || b >= B) continue;
        C v[8][3];
        Real vv[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
          const Real* rw = row_at<Real>(d, r.idx[x], b);
          const C* rc = reinterpret_cast<const C*>(rw) + 3u * a;
          const C e = x < (int)r.nslot ? fac[r.trow + (uint32_t)x * B + b] : C{(Real)1, (Real)0};
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const C el = rc[i];
            phase_elem<Real>(e, el.x, el.y, &v[x][i].x, &v[x][i].y);
          }
          vv[x] = a < sv ? rw[2 * sc + a] : (Real)0;
        }
        C acc[3] = {Elem<C>::zero(), Elem<C>::zero(), Elem<C>::zero()};
        Real va = 0;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
          const Real wx = r.w[x];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            acc[i].x = fma_r(wx, v[x][i].x, acc[i].x);
            acc[i].y = fma_r(wx, v[x][i].y, acc[i].y);
          }
          va = fma_r(wx, vv[x], va);
        }
        if (a < sv) __builtin_nontemporal_store(va, vals + row * sv + (uint64_t)b * sv + a);
        // rip_gamma_complex  src/interpolator_gamma.tpp:100-114, atom by atom
        C* dst = vecs + row * sc + (uint64_t)b * sc + 3u * dsta;
        typedef Real real2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          C tt = Elem<C>::zero();
#pragma unroll
          for (int j = 0; j < 3; ++j) tt = Elem<C>::add(tt, Elem<C>::scale(acc[j], R[3 * i + j]));
          const C o = cmul(ph, tt);
          __builtin_nontemporal_store(real2{o.x, o.y}, reinterpret_cast<real2*>(dst + i));
        }
      }
      return;
    }
    if (!node_order) {
      // Point-major: wave w takes units w, w + TB/64, ... of (point, mode),
      // so it stays on a point for its share of the modes and keeps the
      // point's rotation data (a second mode in flight per wave doubled the
      // registers and made the 64-grid batch of 1e7 points slower, 466 -> 511 ms)
      uint32_t last = 0xFFFFFFFFu;
      for (uint32_t c = wave; c < units; c += TB / 64) {
        const uint32_t pl = c / B, b = c - pl * B;
        const PointRec<Real>& r = s_rec[pl];
        const uint64_t row = ((out_row ? (uint64_t)out_row[pl] : row0 + pl)) * B;
        Real* ov = vals + row * d.sv;
        C* oc = vecs + row * d.sc;
        if (r.nslot == 0) {
          for (uint32_t e = lane; e < d.sv; e += 64) ov[(uint64_t)b * d.sv + e] = 0;
          for (uint32_t e = lane; e < d.sc; e += 64) oc[(uint64_t)b * d.sc + e] = Elem<C>::zero();
          continue;
        }
        if (two) unit_wide<Real, true>(g, d, r, s_q + 3 * pl, b, ov, oc, lds, gamma, cart, wc, pl != last, tab);
        else unit_wide<Real, false>(g, d, r, s_q + 3 * pl, b, ov, oc, lds, gamma, cart, wc, pl != last, tab);
        last = pl;
      }
      return;
    }
    // Mode-major (node-ordered tiles): wave w takes modes w, w + TB/64, ...
    // and for each walks the tile's points in their node order.  Consecutive
    // points of one table block (cube or tetrahedron, same trow) have the
    // same slot rows and factors, which stay in the wave's registers; they
    // are loaded again only when the block changes.  Wave-uniform loops.
    // With `wg`, the points go in chunks whose Gamma phases the block first
    // computes into LDS (once per point and atom instead of once per mode),
    // and the rotation rows and atom map come from per-operation LDS tables.
    // The block sums modes b_lo .. b_hi - 1 only.
    const bool pre = gamma && wg.ph != nullptr;
    const uint32_t na = d.sc / 3u;
    const uint32_t chunk = pre ? (wg.cap / na < npts ? wg.cap / na : npts) : npts;
    const uint32_t b_end = b_hi < B ? b_hi : B;
    // Atom-major (Gamma rows with the factor table and at most one value per
    // atom): lane l takes atom a = l % na of mode l / na, so a wave sums
    // 64 / na modes of a point at once, each lane its atom's three elements:
    // the rotation needs no exchange between lanes.  Same per-element
    // arithmetic as the wave-per-unit walk below, so the same bits.  Points
    // without a table block are left to that walk.
    // (Staging a wave's rotated rows in LDS, to write them as one contiguous
    // run per point instead of three 8-byte stores per lane at a permuted
    // atom, saved no time: 64-grid batch of 1e7 points 140 -> 145 ms.)
    // kWideAtom: the caller chose it for such rows; kWideRest: no point here
    // has a table block
    const bool by_atom = SC == kWideAtom || (SC == kWideAll && pre && tab && d.sv <= na);
    auto atom_walk = [&](uint32_t c0, uint32_t c1) {
      const uint32_t mpw = 64u / na;  // modes per wave pass
      const uint32_t m = lane / na, a = lane - m * na;
      const uint32_t nmodes = b_end > b_lo ? b_end - b_lo : 0u;
      const uint32_t passes = (nmodes + mpw - 1) / mpw;
      const C* fac = wide_factors<Real>(d);
      const uint32_t sc = d.sc, sv = d.sv;
      // the points this walk takes: located, with a table block
      auto taken = [&](uint32_t pl) { return s_rec[pl].nslot != 0 && s_rec[pl].trow != kNoCube; };
      for (uint32_t t = wave; t < passes; t += TB / 64) {
        const uint32_t b = b_lo + t * mpw + m;
        const bool on = m < mpw && b < b_end;
        // Runs of consecutive points of one table block: the block's rows of
        // this lane's mode and atom are loaded and antiphased once per run,
        // and stay loop-invariant over the run's points.
        uint32_t p0 = c0;
        while (true) {
          while (p0 < c1 && !taken(p0)) {
            // without the wave-per-unit walk the unlocated points are zeroed here
            if (SC == kWideAtom && s_rec[p0].nslot == 0 && on) {
              const uint64_t row = ((out_row ? (uint64_t)out_row[p0] : row0 + p0)) * B;
              C* oc = vecs + row * sc + (uint64_t)b * sc;
#pragma unroll
              for (int i = 0; i < 3; ++i) oc[3u * a + i] = Elem<C>::zero();
              if (a < sv) vals[row * sv + (uint64_t)b * sv + a] = 0;
            }
            ++p0;
          }
          if (p0 >= c1) break;
          const uint32_t trow = s_rec[p0].trow;
          uint32_t p1 = p0 + 1;
          while (p1 < c1 && taken(p1) && s_rec[p1].trow == trow) ++p1;
          C v[8][3];
          Real vv[8];
          {
            const PointRec<Real>& rb = s_rec[p0];
#pragma unroll
            for (int x = 0; x < 8; ++x) {
              if (on) {
                const Real* rw = row_at<Real>(d, rb.idx[x], b);
                const C* rc = reinterpret_cast<const C*>(rw) + 3u * a;
                const C e = x < (int)rb.nslot ? fac[trow + (uint32_t)x * B + b] : C{(Real)1, (Real)0};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                  const C el = rc[i];
                  phase_elem<Real>(e, el.x, el.y, &v[x][i].x, &v[x][i].y);
                }
                vv[x] = a < sv ? rw[2 * sc + a] : (Real)0;
              }
            }
          }
          for (uint32_t pl = p0; pl < p1; ++pl) {
            if (!on) continue;
            const PointRec<Real>& r = s_rec[pl];
            const uint64_t row = ((out_row ? (uint64_t)out_row[pl] : row0 + pl)) * B;
            Real* ov = vals + row * sv + (uint64_t)b * sv;
            C* oc = vecs + row * sc + (uint64_t)b * sc;
            C acc[3] = {Elem<C>::zero(), Elem<C>::zero(), Elem<C>::zero()};
            Real va = 0;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
              const Real wx = r.w[x];
#pragma unroll
              for (int i = 0; i < 3; ++i) {
                acc[i].x = fma_r(wx, v[x][i].x, acc[i].x);
                acc[i].y = fma_r(wx, v[x][i].y, acc[i].y);
              }
              va = fma_r(wx, vv[x], va);
            }
            if (a < sv) __builtin_nontemporal_store(va, ov + a);
            // rip_gamma_complex  src/interpolator_gamma.tpp:100-114, atom by atom
            const Real* R = wg.rot + 9u * r.invrot;
            const C ph = wg.ph[(pl - c0) * na + a];
            C* dst = oc + 3u * wg.dst[r.invrot * na + a];
            typedef Real real2 __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              C tt = Elem<C>::zero();
#pragma unroll
              for (int j = 0; j < 3; ++j) tt = Elem<C>::add(tt, Elem<C>::scale(acc[j], R[3 * i + j]));
              const C o = cmul(ph, tt);
              __builtin_nontemporal_store(real2{o.x, o.y}, reinterpret_cast<real2*>(dst + i));
            }
          }
          p0 = p1;
        }
      }
    };
    // Wave per (point, mode): wave w takes modes w, w + TB/64, ... of the
    // block's group; with by_atom only the points the atom walk leaves.
    auto walk = [&](auto two_tag, uint32_t c0, uint32_t c1, bool untabled_only) {
      constexpr bool TWO = decltype(two_tag)::value;
      WideRows<Real, TWO> w;
      for (uint32_t b = b_lo + wave; b < b_end; b += TB / 64) {
        uint32_t cached = kNoCube;
        for (uint32_t pl = c0; pl < c1; ++pl) {
          const PointRec<Real>& r = s_rec[pl];
          if (untabled_only && r.nslot != 0 && r.trow != kNoCube) continue;
          const uint64_t row = ((out_row ? (uint64_t)out_row[pl] : row0 + pl)) * B;
          Real* ov = vals + row * d.sv;
          C* oc = vecs + row * d.sc;
          if (r.nslot == 0) {
            for (uint32_t e = lane; e < d.sv; e += 64) ov[(uint64_t)b * d.sv + e] = 0;
            for (uint32_t e = lane; e < d.sc; e += 64) oc[(uint64_t)b * d.sc + e] = Elem<C>::zero();
            continue;
          }
          const bool reuse = tab && r.trow != kNoCube && r.trow == cached;
          if (!reuse) {
            wide_load<Real, TWO>(d, r, b, tab, w);
            wide_phase_rows<Real, TWO>(w);
            cached = (tab && r.trow != kNoCube) ? r.trow : kNoCube;
          }
          if (pre) {
#pragma unroll
            for (int h = 0; h < (TWO ? 2 : 1); ++h) {
              const uint32_t el = lane + 64u * h;
              if (el < d.sc) {
                const uint32_t a = el / 3u, i = el - 3u * a;
                wc.ph[h] = wg.ph[(pl - c0) * na + a];
#pragma unroll
                for (int j = 0; j < 3; ++j) wc.R[h][j] = wg.rot[9u * r.invrot + 3u * i + j];
                wc.dst[h] = 3u * wg.dst[r.invrot * na + a] + i;
              }
            }
          }
          wide_sum_store<Real, TWO, true>(g, d, r, s_q + 3 * pl, b, w, ov, oc, lds, gamma, cart, wc, !pre);
        }
      }
    };
    for (uint32_t c0 = 0; c0 < npts; c0 += chunk) {
      const uint32_t c1 = c0 + chunk < npts ? c0 + chunk : npts;
      if (pre) {
        __syncthreads();  // the previous chunk's phases are read
        for (uint32_t k = threadIdx.x; k < (c1 - c0) * na; k += TB) {
          const uint32_t pl = c0 + k / na, a = k - (k / na) * na;
          const PointRec<Real>& r = s_rec[pl];
          wg.ph[k] = r.nslot ? gamma_phase<Real>(d, s_q + 3 * pl, r.invrot, a) : Elem<C>::zero();
        }
        __syncthreads();
      }
      if (by_atom) atom_walk(c0, c1);
      // (kWideAtom compiles without the wave-per-unit walk: 109 instead of
      // 240 VGPRs; its points without a table block go to the kWideRest
      // launch.  64-grid batch of 1e7 points 103 -> 87-90 ms.)
      if constexpr (SC != kWideAtom) {
        if (two) walk(std::true_type{}, c0, c1, by_atom);
        else walk(std::false_type{}, c0, c1, by_atom);
      }
    }
    return;
  }
  // Gamma frames: per point in s_q, or per point-group operation in op_frames
  const GammaFrame<Real>* frames = reinterpret_cast<const GammaFrame<Real>*>(s_q);
  auto frame_of = [&](uint32_t pl) { return op_frames ? op_frames + s_rec[pl].invrot : frames + pl; };
  uint32_t u0 = 0;
  constexpr bool kCoop = (L == LAYOUT_GAMMA || L == LAYOUT_SCALAR) && coop_lanes<Real, SC, SV>() > 1;
  if constexpr (kCoop) {
    constexpr uint32_t G = coop_lanes<Real, SC, SV>();
    if (table) {
      u0 = n_coop * B;
      const uint32_t q = threadIdx.x & (G - 1u);
      const uint32_t rb = payload_stride<Real>(d) * (uint32_t)sizeof(Real);  // bytes per row
      const uint32_t slot_bytes = B * rb;
      constexpr uint32_t NG = TB / G;  // lane groups per block
      Chunk<Real> rows[8];
      if (B <= NG) {
        // Group grp keeps mode b = grp % B and walks the points grp / B, + P,
        // + 2P, ...: with the points in node order it meets the same cube or
        // tetrahedron again and again, and reloads the rows only when the
        // block changes.
        const uint32_t P = NG / B;
        const uint32_t grp = threadIdx.x / G;
        const uint32_t b = grp % B;
        uint32_t cached = kNoCube;
        for (uint32_t pl = grp / B; pl < n_coop && grp < P * B; pl += P) {
          const PointRec<Real>& r = s_rec[pl];
          if (r.trow != cached) {
            coop_load<Real>(d, r, (r.trow + b) * rb + q * 16u, slot_bytes, rows);
            cached = r.trow;
          }
          const uint64_t row = ((out_row ? (uint64_t)out_row[pl] : row0 + pl)) * B + b;
          unit_coop<Real, SC, SV, L == LAYOUT_GAMMA, (int)G>(r, rows, q, vals + row * d.sv, vecs + row * d.sc,
                                                             s_ph + NPH * pl, frame_of(pl));
        }
      } else {
        constexpr uint32_t STEP = NG;  // units per block and iteration
        // (point, mode) of unit u, advanced by STEP units per iteration without a division
        uint32_t u = threadIdx.x / G;
        uint32_t pl = u / B, b = u - pl * B;
        const uint32_t sp = STEP / B, sb = STEP - sp * B;
        for (; u < u0; u += STEP) {
          const PointRec<Real>& r = s_rec[pl];
          const uint64_t row = ((out_row ? (uint64_t)out_row[pl] : row0 + pl)) * B + b;
          coop_load<Real>(d, r, (r.trow + b) * rb + q * 16u, slot_bytes, rows);
          unit_coop<Real, SC, SV, L == LAYOUT_GAMMA, (int)G>(r, rows, q, vals + row * d.sv, vecs + row * d.sc,
                                                             s_ph + NPH * pl, frame_of(pl));
          b += sb;
          pl += sp;
          if (b >= B) {
            b -= B;
            ++pl;
          }
        }
      }
    }
  }
  for (uint32_t u = u0 + threadIdx.x; u < units; u += TB) {
    const uint32_t pl = u / B, b = u - pl * B;
    const PointRec<Real>& r = s_rec[pl];
    const uint64_t row = ((out_row ? (uint64_t)out_row[pl] : row0 + pl)) * B + b;
    Real* ov = vals + row * d.sv;
    C* oc = vecs + row * d.sc;
    if (r.nslot == 0) {
      // failed points: zero rows (the call reports the error)
      for (uint32_t e = 0; e < d.sv; ++e) ov[e] = 0;
      for (uint32_t e = 0; e < d.sc; ++e) oc[e] = Elem<C>::zero();
      continue;
    }
    // with the cooperative sums, every table point went there
    if constexpr (L == LAYOUT_GAMMA) {
      unit_onepass<Real, SC, SV, true, GRP>(d, r, b, ov, oc, s_ph + NPH * pl, frame_of(pl), table && !kCoop);
    } else if constexpr (L == LAYOUT_SCALAR) {
      unit_onepass<Real, SC, SV, false, GRP>(d, r, b, ov, oc, nullptr, nullptr, table && !kCoop);
    } else {
      unit_generic<Real, !Fast>(g, d, r, s_q + 3 * pl, b, ov, oc, ph_in_lds ? s_ph + NPH * pl : nullptr);
    }
  }
}

__device__ __forceinline__ void report_point(ErrRec* err, uint64_t index, int status) {
  atomicMin(&err->first_bad, err_key(index, status));
  atomicAdd(&err->count, 1ull);
}

// ---------------------------------------------------------------------------
// the fused kernel: one tile of TILE input points, both phases.  Plain
// (Fast = false) is the literal restatement: exact fold-in and location,
// atan2/sincos antiphases.

template <typename Real, bool Fast, int L, int SC, int SV>
__global__ __launch_bounds__(TILE) void k_ir_interpolate(GridConst g, DataConst d, LaunchArgs a) {
  using C = typename Cplx<Real>::T;
  using S = TileShape<L, SC, TILE, Real>;
  __shared__ PointRec<Real> s_rec[TILE];
  __shared__ C s_ph[TILE * S::NPH];
  __shared__ double s_q[S::NQ];
  __shared__ C s_wide[S::NW];
  const uint64_t tile0 = (uint64_t)blockIdx.x * TILE;
  const uint32_t tid = threadIdx.x;
  const uint64_t p = tile0 + tid;
  const bool want_phase = d.vec_rot == BRILLE_ROT_GAMMA && d.ec[1] > 0;
  const uint32_t n_atoms = d.ec[1] / 3u;
  const bool ph_in_lds = L == LAYOUT_GAMMA || (want_phase && n_atoms <= 4);
  {
    PointRec<Real>& r = s_rec[tid];
    r.nslot = 0;
    const bool live = p < a.n;
    Located P;
    P.status = BRILLE_PT_OK;
    if (live) {
      double Qp[3] = {a.Q[3 * p], a.Q[3 * p + 1], a.Q[3 * p + 2]};
      fold_node<Fast>(g, Qp, a.fold, P);
    }
    if (live && P.status == BRILLE_PT_OK) node_slots(g, P);
    if (live) {
      write_aux(a.aux, p, P);
      if (P.status != BRILLE_PT_OK) report_point(a.err, a.base + p, P.status);
      else store_record<Real, L, SC>(g, d, P, r, s_ph, s_q, tid, ph_in_lds, n_atoms);
    }
  }
  __syncthreads();
  const uint64_t left = a.n - tile0;
  const uint32_t npts = left < (uint64_t)TILE ? (uint32_t)left : (uint32_t)TILE;
  tile_sums<Real, Fast, L, SC, SV>(g, d, s_rec, s_ph, s_q, nullptr, tile0, npts, ph_in_lds, a.vals, a.vecs, s_wide);
}

}  // namespace brille_amd
