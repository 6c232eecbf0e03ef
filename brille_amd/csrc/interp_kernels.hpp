// Fused ir_interpolate_at tile kernel.
//
// One workgroup owns a tile of TILE points.
//   phase 1  (one lane per point): fold-in (stage a) and point location
//            (stage b) in FP64; the (index, weight) pairs, Ridx/invRidx and
//            the irreducible q go to LDS.
//   phase 2  (one lane per (point, mode) unit, consecutive lanes take
//            consecutive modes of a point so loads of vertex rows and stores
//            of output rows coalesce): weighted sums of values and of
//            antiphased vectors (stage c1) fused with the symmetry rotation
//            (stage c2), written straight to the output rows.
// Real = double reproduces the reference's FP64 arithmetic; Real = float runs
// the sums and the rotation in FP32/complex64 with the same FP64 indices and
// weights.
#pragma once
#include "exact_path.hpp"
#include "fast_fold.hpp"

namespace brille_amd {

constexpr int TILE = 256;

template <typename Real>
struct Cplx;
template <>
struct Cplx<double> {
  using T = double2;
};
template <>
struct Cplx<float> {
  using T = float2;
};

struct ErrRec {
  unsigned long long first_bad;  // (index << 3) | point status, min over failing points
  unsigned long long count;
};

struct LocRec {
  double q[3];
  uint32_t idx[8];
  double w[8];
  uint32_t rot, invrot;
  uint8_t n;
  uint8_t status;
};

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

template <typename Real>
__device__ __forceinline__ Real load_val(const DataConst& d, uint64_t i);
template <>
__device__ __forceinline__ double load_val<double>(const DataConst& d, uint64_t i) { return d.vals64[i]; }
template <>
__device__ __forceinline__ float load_val<float>(const DataConst& d, uint64_t i) { return d.vals32[i]; }

template <typename Real>
__device__ __forceinline__ typename Cplx<Real>::T load_vec(const DataConst& d, uint64_t i);
template <>
__device__ __forceinline__ double2 load_vec<double>(const DataConst& d, uint64_t i) { return d.vecs64[i]; }
template <>
__device__ __forceinline__ float2 load_vec<float>(const DataConst& d, uint64_t i) { return d.vecs32[i]; }

// utils::antiphase  src/utilities.tpp:567-577: e^{-i theta} of the Hermitian
// product of a and b.  FP64 keeps the reference's atan2/cos/sin; FP32 uses
// the identical closed form conj(z)/|z| (cos(atan2(y,x)) = x/|z|,
// sin(atan2(y,x)) = y/|z|) which needs no transcendental.
template <typename Real>
__device__ __forceinline__ typename Cplx<Real>::T antiphase(const DataConst& d, uint64_t a, uint64_t b, uint32_t n);
template <>
__device__ __forceinline__ double2 antiphase<double>(const DataConst& d, uint64_t a, uint64_t b, uint32_t n) {
  double re = 0, im = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double2 x = d.vecs64[a + i], y = d.vecs64[b + i];
    re += x.x * y.x + x.y * y.y;
    im += x.x * y.y - x.y * y.x;
  }
  double th = -1.0 * atan2(im, re);
  double s, c;
  sincos(th, &s, &c);
  return make_double2(1.0 * c, 1.0 * s);
}
template <>
__device__ __forceinline__ float2 antiphase<float>(const DataConst& d, uint64_t a, uint64_t b, uint32_t n) {
  float re = 0.f, im = 0.f;
  for (uint32_t i = 0; i < n; ++i) {
    float2 x = d.vecs32[a + i], y = d.vecs32[b + i];
    re += x.x * y.x + x.y * y.y;
    im += x.x * y.y - x.y * y.x;
  }
  float r2 = re * re + im * im;
  if (!(r2 > 0.f)) return make_float2(re < 0.f ? -1.f : 1.f, 0.f);
  float inv = rsqrtf(r2);
  return make_float2(re * inv, -im * inv);
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

// Phase 2 for one (point, mode) unit: values.
template <typename Real>
__device__ inline void unit_values(const GridConst& g, const DataConst& d, const LocRec& r, uint32_t b, Real* out) {
  const uint32_t sv = d.sv, ns = d.ev[0], nv = d.ev[1] / 3u, nm = d.ev[2] / 9u;
  const uint32_t pivot = r.idx[0];
  uint64_t base[8];
  Real w[8];
  for (int x = 0; x < r.n; ++x) {
    base[x] = ((uint64_t)r.idx[x] * d.B + perm_at(d, pivot, r.idx[x], b)) * sv;
    w[x] = (Real)r.w[x];
  }
  Real* o = out + (uint64_t)b * sv;
  uint32_t e = 0;
  for (; e < ns; ++e) {
    Real acc = 0;
    for (int x = 0; x < r.n; ++x) acc += w[x] * load_val<Real>(d, base[x] + e);
    o[e] = acc;
  }
  for (uint32_t k = 0; k < nv; ++k, e += 3) {
    Real v[3] = {0, 0, 0};
    for (int x = 0; x < r.n; ++x)
#pragma unroll
      for (int j = 0; j < 3; ++j) v[j] += w[x] * load_val<Real>(d, base[x] + e + j);
    rotate_vec3<Real, Real>(d.val_rot, d.val_lu, g, r.rot, r.invrot, v);
    o[e] = v[0]; o[e + 1] = v[1]; o[e + 2] = v[2];
  }
  for (uint32_t k = 0; k < nm; ++k, e += 9) {
    Real m[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) m[j] = 0;
    for (int x = 0; x < r.n; ++x)
#pragma unroll
      for (int j = 0; j < 9; ++j) m[j] += w[x] * load_val<Real>(d, base[x] + e + j);
    rotate_mat3<Real, Real>(d.val_rot, d.val_lu, g, r.rot, r.invrot, m);
#pragma unroll
    for (int j = 0; j < 9; ++j) o[e + j] = m[j];
  }
}

// e_iqd for one atom of one point  src/interpolator_gamma.tpp:18-32
template <typename Real>
__device__ inline typename Cplx<Real>::T gamma_phase_one(const DataConst& d, const LocRec& r, uint32_t k) {
  const uint32_t key = k * d.gamma_ops + r.invrot;
  const double* dv = d.gvec + 3 * (uint64_t)d.vidx[key];
  double dot = 0;
  dot += r.q[0] * dv[0];
  dot += r.q[1] * dv[1];
  dot += r.q[2] * dv[2];
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

// Phase 2 for one (point, mode) unit: antiphased vectors + rotation.
template <typename Real>
__device__ inline void unit_vectors(const GridConst& g, const DataConst& d, const LocRec& r, uint32_t b,
                                    typename Cplx<Real>::T* out, const typename Cplx<Real>::T* phases) {
  using C = typename Cplx<Real>::T;
  const uint32_t sc = d.sc, ns = d.ec[0], nv = d.ec[1] / 3u, nm = d.ec[2] / 9u;
  const uint32_t pivot = r.idx[0];
  uint64_t base[8];
  C cw[8];
  const uint64_t pbase = ((uint64_t)pivot * d.B + b) * sc;
  for (int x = 0; x < r.n; ++x) {
    base[x] = ((uint64_t)r.idx[x] * d.B + perm_at(d, pivot, r.idx[x], b)) * sc;
    C e = antiphase<Real>(d, pbase, base[x], sc);
    Real w = (Real)r.w[x];
    cw[x].x = w * e.x;  // weights[x]*eith
    cw[x].y = w * e.y;
  }
  C* o = out + (uint64_t)b * sc;
  uint32_t e = 0;
  for (; e < ns; ++e) {
    C acc = Elem<C>::zero();
    for (int x = 0; x < r.n; ++x) acc = Elem<C>::add(acc, cmul(cw[x], load_vec<Real>(d, base[x] + e)));
    o[e] = acc;
  }
  const bool gamma = d.vec_rot == BRILLE_ROT_GAMMA;
  const bool cart = d.vec_lu == BRILLE_LU_ANGSTROM;
  for (uint32_t k = 0; k < nv; ++k, e += 3) {
    C v[3] = {Elem<C>::zero(), Elem<C>::zero(), Elem<C>::zero()};
    for (int x = 0; x < r.n; ++x)
#pragma unroll
      for (int j = 0; j < 3; ++j) v[j] = Elem<C>::add(v[j], cmul(cw[x], load_vec<Real>(d, base[x] + e + j)));
    if (gamma) {
      // rip_gamma_complex  src/interpolator_gamma.tpp:100-114
      C t[3];
      if (cart) rot3<C, double, Real>(t, d.cart_rot + 9 * (uint64_t)r.invrot, v);
      else rot3<C, int32_t, Real>(t, g.rot + 9 * (uint64_t)r.invrot, v);
      const uint32_t key = k * d.gamma_ops + r.invrot;
      const uint32_t l = d.F0[key];
      C ph = phases ? phases[k] : gamma_phase_one<Real>(d, r, k);
      C* dst = o + ns + 3u * l;
#pragma unroll
      for (int j = 0; j < 3; ++j) dst[j] = cmul(ph, t[j]);
    } else {
      rotate_vec3<C, Real>(d.vec_rot, d.vec_lu, g, r.rot, r.invrot, v);
      o[e] = v[0]; o[e + 1] = v[1]; o[e + 2] = v[2];
    }
  }
  for (uint32_t k = 0; k < nm; ++k, e += 9) {
    C m[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) m[j] = Elem<C>::zero();
    for (int x = 0; x < r.n; ++x)
#pragma unroll
      for (int j = 0; j < 9; ++j) m[j] = Elem<C>::add(m[j], cmul(cw[x], load_vec<Real>(d, base[x] + e + j)));
    rotate_mat3<C, Real>(d.vec_rot, d.vec_lu, g, r.rot, r.invrot, m);
#pragma unroll
    for (int j = 0; j < 9; ++j) o[e + j] = m[j];
  }
}

// e_iqd for every atom of one point
template <typename Real>
__device__ inline void gamma_phases(const DataConst& d, const LocRec& r, typename Cplx<Real>::T* ph, uint32_t n_atoms) {
  for (uint32_t k = 0; k < n_atoms; ++k) ph[k] = gamma_phase_one<Real>(d, r, k);
}

// Payload layouts with a specialised phase 2 (the general one handles all).
enum Layout : int {
  LAYOUT_GENERIC = 0,  // any elements, rotations and permutation table
  LAYOUT_GAMMA = 1,    // scalar values; vectors = one 3-vector per atom rotating like Gamma; no permutations
  LAYOUT_SCALAR = 2    // scalar values and scalar vectors; no permutations
};

// Phase 2, specialised: scalar values of one (point, mode) unit.
template <typename Real>
__device__ __forceinline__ void unit_values_scalar(const DataConst& d, const LocRec& r, uint32_t b, Real* out) {
  const uint32_t sv = d.sv;
  Real* o = out + (uint64_t)b * sv;
  for (uint32_t e = 0; e < sv; ++e) {
    Real acc = 0;
    for (int x = 0; x < r.n; ++x) acc += (Real)r.w[x] * load_val<Real>(d, ((uint64_t)r.idx[x] * d.B + b) * sv + e);
    o[e] = acc;
  }
}

// Phase 2, specialised: antiphased Gamma-rotated eigenvectors of one (point, mode) unit.
template <typename Real>
__device__ __forceinline__ void unit_vectors_gamma(const GridConst& g, const DataConst& d, const LocRec& r, uint32_t b,
                                                   typename Cplx<Real>::T* out, const typename Cplx<Real>::T* phases,
                                                   bool cart) {
  using C = typename Cplx<Real>::T;
  const uint32_t sc = d.sc, nv = sc / 3u;
  const uint64_t pb = ((uint64_t)r.idx[0] * d.B + b) * sc;
  C cw[8];
  const int n = r.n;
  for (int x = 0; x < n; ++x) {
    C e;
    if (x == 0) {
      // the pivot against itself: a positive real Hermitian product, e^{-i0} = 1
      e.x = 1;
      e.y = 0;
    } else {
      e = antiphase<Real>(d, pb, ((uint64_t)r.idx[x] * d.B + b) * sc, sc);
    }
    Real w = (Real)r.w[x];
    cw[x].x = w * e.x;
    cw[x].y = w * e.y;
  }
  C* o = out + (uint64_t)b * sc;
  const int32_t* iR = g.rot + 9 * (uint64_t)r.invrot;
  const double* cR = cart ? d.cart_rot + 9 * (uint64_t)r.invrot : nullptr;
  for (uint32_t k = 0; k < nv; ++k) {
    C v[3] = {Elem<C>::zero(), Elem<C>::zero(), Elem<C>::zero()};
    for (int x = 0; x < n; ++x) {
      const uint64_t base = ((uint64_t)r.idx[x] * d.B + b) * sc + 3u * k;
#pragma unroll
      for (int j = 0; j < 3; ++j) v[j] = Elem<C>::add(v[j], cmul(cw[x], load_vec<Real>(d, base + j)));
    }
    C t[3];
    if (cart) rot3<C, double, Real>(t, cR, v);
    else rot3<C, int32_t, Real>(t, iR, v);
    const uint32_t l = d.F0[k * d.gamma_ops + r.invrot];
    C ph = phases ? phases[k] : gamma_phase_one<Real>(d, r, k);
    C* dst = o + 3u * l;
#pragma unroll
    for (int j = 0; j < 3; ++j) dst[j] = cmul(ph, t[j]);
  }
}

// Phase 2, specialised: scalar (unrotated) vectors of one (point, mode) unit.
template <typename Real>
__device__ __forceinline__ void unit_vectors_scalar(const DataConst& d, const LocRec& r, uint32_t b,
                                                    typename Cplx<Real>::T* out) {
  using C = typename Cplx<Real>::T;
  const uint32_t sc = d.sc;
  const uint64_t pb = ((uint64_t)r.idx[0] * d.B + b) * sc;
  C cw[8];
  for (int x = 0; x < r.n; ++x) {
    C e = antiphase<Real>(d, pb, ((uint64_t)r.idx[x] * d.B + b) * sc, sc);
    Real w = (Real)r.w[x];
    cw[x].x = w * e.x;
    cw[x].y = w * e.y;
  }
  C* o = out + (uint64_t)b * sc;
  for (uint32_t e = 0; e < sc; ++e) {
    C acc = Elem<C>::zero();
    for (int x = 0; x < r.n; ++x) acc = Elem<C>::add(acc, cmul(cw[x], load_vec<Real>(d, ((uint64_t)r.idx[x] * d.B + b) * sc + e)));
    o[e] = acc;
  }
}

struct LaunchArgs {
  const double* Q;
  uint64_t n;
  uint64_t base;  // index of Q[0] in the caller's full array (for error reporting)
  void* vals;
  void* vecs;
  brille_aux_out aux;
  ErrRec* err;
  int no_move;
  int use_proposal;
};

// phase 1 for one point
template <bool Fast>
__device__ inline void locate_point(const GridConst& g, const double* Q, bool no_move, LocRec& r, uint32_t* node_out,
                                    uint32_t* tet_out, int32_t* tau_out) {
  const Tol tbz{g.tol_bz_rel, g.tol_bz_abs};
  const Tol tdef{g.tol_def_rel, g.tol_def_abs};
  int32_t tau[3] = {0, 0, 0};
  r.n = 0;
  r.status = BRILLE_PT_OK;
  r.rot = r.invrot = g.identity;
  *node_out = 0xFFFFFFFFu;
  *tet_out = kNoTet;
  if (no_move) {
    r.q[0] = Q[0]; r.q[1] = Q[1]; r.q[2] = Q[2];
  } else {
    int st = Fast ? fast_moveinto(g, Q, r.q, tau, tbz) : moveinto_exact(g, Q, r.q, tau, tbz);
    if (st == BRILLE_PT_OK) st = Fast ? fast_wedge_search(g, r.q, &r.rot, &r.invrot, tbz, tdef)
                                      : wedge_search_exact(g, r.q, &r.rot, &r.invrot, tbz, tdef);
    if (st != BRILLE_PT_OK) { r.status = (uint8_t)st; tau_out[0] = tau[0]; tau_out[1] = tau[1]; tau_out[2] = tau[2]; return; }
  }
  tau_out[0] = tau[0]; tau_out[1] = tau[1]; tau_out[2] = tau[2];
  double x[3];
  mat_vec(x, g.outer_to_xyz, r.q);  // brille::get_xyz
  uint64_t node = 0;
  if (!node_subscript<Fast>(g, x, &node, tdef)) { r.status = BRILLE_PT_NULL_NODE; return; }
  *node_out = (uint32_t)node;
  int nw;
  if (g.node_type[node] == BRILLE_NODE_CUBE) nw = cube_weights(g, g.node_index[node], x, r.idx, r.w, tdef);
  else nw = poly_weights(g, g.node_index[node], x, r.idx, r.w, tet_out, tdef);
  r.n = (uint8_t)nw;
  if (nw < 1) r.status = BRILLE_PT_NOT_FOUND;
}

template <typename Real, bool Fast, int L>
__global__ __launch_bounds__(TILE) void k_ir_interpolate(GridConst g, DataConst d, LaunchArgs a) {
  using C = typename Cplx<Real>::T;
  __shared__ LocRec s_rec[TILE];
  __shared__ C s_ph[TILE * 4];
  const uint64_t tile0 = (uint64_t)blockIdx.x * TILE;
  const uint32_t tid = threadIdx.x;
  const uint64_t p = tile0 + tid;
  const bool want_phase = d.vec_rot == BRILLE_ROT_GAMMA && d.ec[1] > 0;
  const uint32_t n_atoms = d.ec[1] / 3u;
  const bool ph_in_lds = want_phase && n_atoms <= 4;
  if (p < a.n) {
    LocRec& r = s_rec[tid];
    double Qp[3] = {a.Q[3 * p], a.Q[3 * p + 1], a.Q[3 * p + 2]};
    uint32_t node, tet;
    int32_t tau[3];
    locate_point<Fast>(g, Qp, a.no_move, r, &node, &tet, tau);
    if (a.aux.ir_q) { a.aux.ir_q[3 * p] = r.q[0]; a.aux.ir_q[3 * p + 1] = r.q[1]; a.aux.ir_q[3 * p + 2] = r.q[2]; }
    if (a.aux.tau) { a.aux.tau[3 * p] = tau[0]; a.aux.tau[3 * p + 1] = tau[1]; a.aux.tau[3 * p + 2] = tau[2]; }
    if (a.aux.rot) a.aux.rot[p] = r.rot;
    if (a.aux.invrot) a.aux.invrot[p] = r.invrot;
    if (a.aux.node) a.aux.node[p] = node;
    if (a.aux.tet) a.aux.tet[p] = tet;
    if (a.aux.n_iw) a.aux.n_iw[p] = r.n;
    if (a.aux.iw_index)
      for (int k = 0; k < 8; ++k) a.aux.iw_index[8 * p + k] = k < r.n ? r.idx[k] : 0u;
    if (a.aux.iw_weight)
      for (int k = 0; k < 8; ++k) a.aux.iw_weight[8 * p + k] = k < r.n ? r.w[k] : 0.0;
    if (a.aux.status) a.aux.status[p] = r.status;
    if (r.status != BRILLE_PT_OK) {
      r.n = 0;
      unsigned long long code = ((unsigned long long)(a.base + p) << 3) | r.status;
      atomicMin(&a.err->first_bad, code);
      atomicAdd(&a.err->count, 1ull);
    } else if (ph_in_lds) {
      gamma_phases<Real>(d, r, s_ph + 4 * tid, n_atoms);
    }
  } else {
    s_rec[tid].n = 0;
  }
  __syncthreads();
  const uint64_t left = a.n - tile0;
  const uint32_t npts = left < (uint64_t)TILE ? (uint32_t)left : (uint32_t)TILE;
  const uint32_t B = d.B;
  const uint32_t units = npts * B;
  Real* vals = (Real*)a.vals;
  C* vecs = (C*)a.vecs;
  for (uint32_t u = tid; u < units; u += TILE) {
    const uint32_t pl = u / B, b = u - pl * B;
    const LocRec& r = s_rec[pl];
    const uint64_t pt = tile0 + pl;
    if (r.n == 0) {
      // failed points: zero rows (the call reports the error)
      Real* ov = vals + (pt * B + b) * d.sv;
      for (uint32_t e = 0; e < d.sv; ++e) ov[e] = 0;
      C* oc = vecs + (pt * B + b) * d.sc;
      for (uint32_t e = 0; e < d.sc; ++e) oc[e] = Elem<C>::zero();
      continue;
    }
    const C* ph = (want_phase && ph_in_lds) ? s_ph + 4 * pl : nullptr;
    if constexpr (L == LAYOUT_GAMMA) {
      unit_values_scalar<Real>(d, r, b, vals + pt * B * d.sv);
      unit_vectors_gamma<Real>(g, d, r, b, vecs + pt * B * d.sc, ph, d.vec_lu == BRILLE_LU_ANGSTROM);
    } else if constexpr (L == LAYOUT_SCALAR) {
      unit_values_scalar<Real>(d, r, b, vals + pt * B * d.sv);
      unit_vectors_scalar<Real>(d, r, b, vecs + pt * B * d.sc);
    } else {
      unit_values<Real>(g, d, r, b, vals + pt * B * d.sv);
      unit_vectors<Real>(g, d, r, b, vecs + pt * B * d.sc, ph);
    }
  }
}

}  // namespace brille_amd
