// Mode sorting on the device: the permutation table of DualInterpolator::sort
// (src/interpolatordual.hpp:398-434) built from the filled payload.
//
// 1. Pairs: every ordered pair of distinct vertices that share a cube node or a
//    tetrahedron (PolyTrellis::collect_keys, src/trellis_poly_parallel.tpp:487-524),
//    as per-vertex neighbour lists: count, scan, place, then sort and
//    de-duplicate each list (a vertex has a few dozen entries).  The keys
//    i * V + j come out sorted, i major.
// 2. One wave per pair i < j: the B x B cost matrix of the pair
//    (DualInterpolator::cost_matrix, src/interpolatordual.hpp:384-396, with
//    Interpolator::add_cost, src/interpolator_cost.tpp:21-58), one entry per
//    lane at a time in the reference's element order, then the exact linear
//    assignment by successive shortest augmenting paths with the columns
//    spread over the lanes (the LDS holds the potentials, the Dijkstra
//    distances and the matching).  The assignment goes to key (i, j), its
//    inverse to (j, i) (src/interpolatordual.hpp:422-432); identity
//    assignments share permutation 0.
#include <algorithm>
#include <vector>

#include "grid_state.hpp"

namespace brille_amd {

namespace {

constexpr int SORT_BLOCK = 64;  // one wave per pair

// ---- pairs ----------------------------------------------------------------

// thread t: position t % 8 of cube t / 8, then position of a tetrahedron;
// PASS 0 counts each ordered pair at its first vertex, PASS 1 places the
// second vertex in the first's list (cursor = the count array, cleared)
template <int PASS>
__global__ __launch_bounds__(256) void k_pair_emit(const uint32_t* cube_v, uint64_t n_cubes, const uint32_t* tet_v,
                                                   uint64_t n_tets, uint32_t* __restrict__ cnt,
                                                   const uint32_t* __restrict__ off, uint32_t* __restrict__ adj) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nc = 8 * n_cubes;
  const uint32_t* u;
  uint32_t k, a;
  if (t < nc) {
    u = cube_v + 8 * (t >> 3);
    k = 8;
    a = (uint32_t)(t & 7u);
  } else if (t < nc + 4 * n_tets) {
    u = tet_v + 4 * ((t - nc) >> 2);
    k = 4;
    a = (uint32_t)((t - nc) & 3u);
  } else {
    return;
  }
  const uint32_t va = u[a];
  for (uint32_t b = 0; b < k; ++b) {
    const uint32_t vb = u[b];
    if (b == a || vb == va) continue;
    if (PASS == 0) atomicAdd(&cnt[va], 1u);
    else adj[off[va] + atomicAdd(&cnt[va], 1u)] = vb;
  }
}

// each vertex's list sorted and de-duplicated in place; uniq[v] = its length
__global__ __launch_bounds__(256) void k_pair_unique(uint64_t V, const uint32_t* __restrict__ off,
                                                     const uint32_t* __restrict__ cnt, uint32_t* __restrict__ adj,
                                                     uint32_t* __restrict__ uniq) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  uint32_t* a = adj + off[v];
  const uint32_t n = cnt[v];
  for (uint32_t i = 1; i < n; ++i) {
    const uint32_t x = a[i];
    uint32_t j = i;
    while (j > 0 && a[j - 1] > x) {
      a[j] = a[j - 1];
      --j;
    }
    a[j] = x;
  }
  uint32_t m = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (m == 0 || a[i] != a[m - 1]) a[m++] = a[i];
  uniq[v] = m;
}

// keys i * V + j in list order, and the neighbour j of each key
__global__ __launch_bounds__(256) void k_pair_keys(uint64_t V, const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ adj, const uint32_t* __restrict__ uniq,
                                                   const uint32_t* __restrict__ koff, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ nbr) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const uint32_t* a = adj + off[v];
  for (uint32_t k = 0; k < uniq[v]; ++k) {
    keys[koff[v] + k] = v * V + a[k];
    nbr[koff[v] + k] = a[k];
  }
}

// rev[k]: the position of key (j, i) for key k = (i, j) (the lists are
// symmetric), by binary search in j's list; one thread per key
__global__ __launch_bounds__(256) void k_pair_rev(uint64_t V, uint64_t n_keys, const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ koff,
                                                  uint32_t* __restrict__ rev) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_keys) return;
  const uint32_t i = (uint32_t)(keys[k] / V), j = nbr[k];
  uint32_t lo = koff[j], hi = koff[j + 1];
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (nbr[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  rev[k] = (lo < koff[j + 1] && nbr[lo] == i) ? lo : 0xFFFFFFFFu;
}

// ---- costs ----------------------------------------------------------------

// approx_float::scalar(|c|, 1) with its default tolerances (relative
// 1e4 x DBL_EPSILON, absolute 5e-15), then the clamp of |c| to one
// (src/utilities.tpp:263-275, 369-373); false when |c| exceeds one by more
__device__ __forceinline__ bool clamp_cos(double& c) {
  double act = fabs(c);
  const double x = fabs(act - 1.0);
  const bool near = x <= 5.0e-15 + (2.220446049250313e-16 * 10000.0) * fabs(act + 1.0) || x < 2.2250738585072014e-308;
  if (near && act > 1) {
    c /= act;
    act = fabs(c);
  }
  return act <= 1;
}

// utils::vector_angle of real vectors  src/utilities.tpp:249-278
template <typename Real>
__device__ double vec_angle_r(uint32_t n, const Real* A, const Real* B, bool& ok) {
  double AA = 0, BB = 0, AB = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const double a = (double)A[i], b = (double)B[i];
    AA += a * a;
    BB += b * b;
    AB += a * b;
  }
  const double nA = sqrt(AA), nB = sqrt(BB);
  double c = (nA != 0 && nB != 0) ? AB / (nA * nB) : ((nA != 0 || nB != 0) ? 0.0 : 1.0);
  if (!clamp_cos(c)) ok = false;
  return acos(c);
}

template <typename Real>
__device__ double vec_dist_r(uint32_t n, const Real* a, const Real* b) {
  double s = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const double d = (double)a[i] - (double)b[i];
    s += d * d;
  }
  return sqrt(s);
}

// the vector cost of real vectors  src/interpolator.hpp:256-293
template <typename Real>
__device__ double vecfun_r(int t, uint32_t n, const Real* a, const Real* b, bool& ok) {
  switch (t) {
    case 1: return vec_dist_r(n, a, b);
    case 2: {
      double h = 0;
      for (uint32_t i = 0; i < n; ++i) h += (double)a[i] * (double)b[i];
      return 1 - h;
    }
    case 3:
    case 4: return vec_angle_r(n, a, b, ok);
    default: {
      const double s = sin(vec_angle_r(n, a, b, ok));
      return s * s;
    }
  }
}

// The second vertex's vector elements after multiplication by the antiphase
// (er, ei): element z of phased = (er + i ei) (br + i bi), as
// utils::inplace_antiphase forms it (src/utilities.tpp:589-593).
template <typename Real>
struct Phased {
  const Real* b;
  double er, ei;
  __device__ double re(uint32_t z) const { return er * (double)b[2 * z] - ei * (double)b[2 * z + 1]; }
  __device__ double im(uint32_t z) const { return er * (double)b[2 * z + 1] + ei * (double)b[2 * z]; }
};

// hermitian_product of a (complex) and phased b from element z0, n elements
template <typename Real>
__device__ void herm_ab(uint32_t z0, uint32_t n, const Real* a, const Phased<Real>& b, double& hr, double& hi) {
  hr = 0;
  hi = 0;
  for (uint32_t i = z0; i < z0 + n; ++i) {
    const double ar = (double)a[2 * i], ai = (double)a[2 * i + 1], br = b.re(i), bi = b.im(i);
    hr += ar * br + ai * bi;
    hi += ar * bi - ai * br;
  }
}

// the vector cost of complex vectors, a against phased b, elements [z0, z0 + n)
// src/interpolator.hpp:256-293 with src/utilities.tpp:280-417
template <typename Real>
__device__ double vecfun_c(int t, uint32_t z0, uint32_t n, const Real* a, const Phased<Real>& b, bool& ok) {
  if (t == 1) {
    double s = 0;
    for (uint32_t i = z0; i < z0 + n; ++i) {
      const double dr = (double)a[2 * i] - b.re(i), di = (double)a[2 * i + 1] - b.im(i);
      s += dr * dr + di * di;
    }
    return sqrt(s);
  }
  if (t == 2) {
    double hr, hi;
    herm_ab(z0, n, a, b, hr, hi);
    return 1 - (hr * hr + hi * hi);
  }
  if (t == 3) {  // euclidean_angle
    double AB = 0, nA = 0, nB = 0;
    for (uint32_t i = z0; i < z0 + n; ++i) {
      const double ar = (double)a[2 * i], ai = (double)a[2 * i + 1], br = b.re(i), bi = b.im(i);
      AB += ar * br + ai * bi;
      nA += ar * ar + ai * ai;
      nB += br * br + bi * bi;
    }
    nA = sqrt(nA);
    nB = sqrt(nB);
    double c = (nA != 0 && nB != 0) ? AB / (nA * nB) : ((nA != 0 || nB != 0) ? 0.0 : 1.0);
    if (!clamp_cos(c)) ok = false;
    return acos(c);
  }
  // hermitian_angle: |a^H b| over the norms
  double hr, hi;
  herm_ab(z0, n, a, b, hr, hi);
  const double nAB = sqrt(hr * hr + hi * hi);
  double aa = 0, bb = 0;
  for (uint32_t i = z0; i < z0 + n; ++i) {
    const double ar = (double)a[2 * i], ai = (double)a[2 * i + 1];
    aa += ar * ar + ai * ai;
  }
  for (uint32_t i = z0; i < z0 + n; ++i) {
    const double br = b.re(i), bi = b.im(i);
    bb += br * br + bi * bi;
  }
  const double nA = sqrt(aa), nB = sqrt(bb);
  double c = (nA != 0 && nB != 0) ? nAB / (nA * nB) : ((nA != 0 || nB != 0) ? 0.0 : 1.0);
  if (!clamp_cos(c)) ok = false;
  const double th = acos(c);
  if (t == 4) return th;
  const double s = sin(th);
  return s * s;
}

struct CostArgs {
  uint32_t B, sv, sc, stride;
  uint32_t ev[3], ec[3];
  int32_t vs, vv, cs, cv;  // scalar / vector cost kinds of values and vectors
  double vw[3], cw[3];
};

// cost of mode a of row block x0 against mode b of row block x1: the values'
// term (no phase) then the vectors' term (after the antiphase), each
// weights[0] s + weights[1] v + weights[2] m, added in that order to zero
// (DualInterpolator::cost_matrix, src/interpolatordual.hpp:384-396)
template <typename Real>
__device__ double pair_cost(const CostArgs& c, const Real* x0, const Real* x1, bool& ok) {
  double cost = 0;
  if (c.sv) {
    const Real* a = x0 + 2 * c.sc;
    const Real* b = x1 + 2 * c.sc;
    const uint32_t mo = c.ev[0] + c.ev[1];
    double s = 0, v = 0, m = 0;
    if (c.ev[0])
      for (uint32_t z = 0; z < c.ev[0]; ++z) s += fabs((double)a[z] - (double)b[z]);
    if (c.ev[1]) v = vecfun_r(c.vv, c.ev[1], a + c.ev[0], b + c.ev[0], ok);
    if (c.ev[2])
      for (uint32_t q = 0; q < c.ev[2] / 9; ++q) m += vec_dist_r(9, a + mo + 9 * q, b + mo + 9 * q);
    cost += c.vw[0] * s + c.vw[1] * v + c.vw[2] * m;
  }
  if (c.sc) {
    // antiphase of the whole span  src/utilities.tpp:567-577
    double rd = 0, id = 0;
    for (uint32_t z = 0; z < c.sc; ++z) {
      const double ar = (double)x0[2 * z], ai = (double)x0[2 * z + 1], br = (double)x1[2 * z], bi = (double)x1[2 * z + 1];
      rd += ar * br + ai * bi;
      id += ar * bi - ai * br;
    }
    const double th = -1.0 * atan2(id, rd);
    const Phased<Real> ph{x1, 1.0 * cos(th), 1.0 * sin(th)};
    const uint32_t mo = c.ec[0] + c.ec[1];
    double s = 0, v = 0, m = 0;
    if (c.ec[0])
      for (uint32_t z = 0; z < c.ec[0]; ++z) {
        const double dr = (double)x0[2 * z] - ph.re(z), di = (double)x0[2 * z + 1] - ph.im(z);
        s += sqrt(dr * dr + di * di);
      }
    if (c.ec[1]) v = vecfun_c(c.cv, c.ec[0], c.ec[1], x0, ph, ok);
    if (c.ec[2])
      for (uint32_t q = 0; q < c.ec[2] / 9; ++q) m += vecfun_c(1, mo + 9 * q, 9, x0, ph, ok);
    cost += c.cw[0] * s + c.cw[1] * v + c.cw[2] * m;
  }
  return cost;
}

// (value, column) pairs: the smaller value, the lower column on ties
__device__ __forceinline__ void wave_argmin(double& v, uint32_t& j) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double ov = __shfl_xor(v, o);
    const uint32_t oj = __shfl_xor(j, o);
    if (ov < v || (ov == v && oj < j)) {
      v = ov;
      j = oj;
    }
  }
}

// forward pairs (i < j) per vertex: count, then (after a scan) the pair's
// vertices and key index, in key order
__global__ __launch_bounds__(256) void k_pair_fwd(uint64_t V, const uint32_t* __restrict__ koff,
                                                  const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ poff,
                                                  uint32_t* __restrict__ cnt, uint32_t* __restrict__ pi,
                                                  uint32_t* __restrict__ pj, uint32_t* __restrict__ pk) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  uint32_t m = 0;
  for (uint32_t k = koff[v]; k < koff[v + 1]; ++k) {
    const uint32_t j = nbr[k];
    if (j <= (uint32_t)v) continue;
    if (poff) {
      pi[poff[v] + m] = (uint32_t)v;
      pj[poff[v] + m] = j;
      pk[poff[v] + m] = k;
    }
    ++m;
  }
  if (!poff) cnt[v] = m;
}

// The cost matrices of pairs [p0, p0 + np): one lane per entry (pair, a, b),
// so every lane of every wave computes (B x B entries per pair).
template <typename Real>
__global__ __launch_bounds__(256) void k_pair_costs(CostArgs c, const Real* rows, const uint32_t* __restrict__ pi,
                                                    const uint32_t* __restrict__ pj, uint64_t p0, uint64_t np,
                                                    double* __restrict__ C, uint32_t* __restrict__ bad) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t B = c.B, BB = B * B;
  if (t >= np * BB) return;
  const uint64_t p = t / BB;
  const uint32_t e = (uint32_t)(t - p * BB);
  const uint32_t a = e / B, b = e - a * B;
  const uint64_t i = pi[p0 + p], j = pj[p0 + p];
  bool ok = true;
  C[t] = pair_cost(c, rows + (i * B + a) * c.stride, rows + (j * B + b) * c.stride, ok);
  if (!ok) atomicOr(bad, 1u);
}

// The easy pairs first, one thread each: when the first minimum of every row
// of the cost matrix lies in a different column, that column-per-row
// assignment is the optimum and is exactly what the augmenting-path solver
// below returns (each row's search then ends at its first minimum, a free
// column).  Such pairs are written here; the others are flagged (hard[q] = 1)
// for the full solver.  Matrices of more than 64 columns all go to the solver.
__global__ __launch_bounds__(256) void k_pair_fast(const double* __restrict__ Cg, uint32_t B, uint64_t p0, uint64_t np,
                                                   const uint32_t* __restrict__ pk, const uint32_t* __restrict__ rev,
                                                   uint32_t* __restrict__ perm_index, uint32_t* __restrict__ perms,
                                                   uint32_t* __restrict__ hard) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  if (B > 64) {
    hard[q] = 1;
    return;
  }
  const double* C = Cg + q * B * B;
  unsigned long long taken = 0;
  bool moved = false, ok = true;
  for (uint32_t a = 0; a < B && ok; ++a) {
    double best = __builtin_huge_val();
    uint32_t bc = 0xFFFFFFFFu;
    for (uint32_t b = 0; b < B; ++b) {
      const double c = C[(uint64_t)a * B + b];
      if (c < best) {
        best = c;
        bc = b;
      }
    }
    if (bc == 0xFFFFFFFFu || (taken >> bc) & 1ull) ok = false;
    else taken |= 1ull << bc;
    moved = moved || bc != a;
  }
  const uint32_t k = pk[p0 + q], rv = rev[k];
  if (!ok || rv == 0xFFFFFFFFu) {
    hard[q] = 1;
    return;
  }
  hard[q] = 0;
  if (!moved) {
    perm_index[k] = 0;
    perm_index[rv] = 0;
    return;
  }
  const uint32_t slot = (uint32_t)(1 + 2 * (p0 + q));
  for (uint32_t a = 0; a < B; ++a) {
    double best = __builtin_huge_val();
    uint32_t bc = 0;
    for (uint32_t b = 0; b < B; ++b) {
      const double c = C[(uint64_t)a * B + b];
      if (c < best) {
        best = c;
        bc = b;
      }
    }
    perms[(uint64_t)slot * B + a] = bc;          // row: column of row a
    perms[(uint64_t)(slot + 1) * B + bc] = a;    // inverse
  }
  perm_index[k] = slot;
  perm_index[rv] = slot + 1;
}

// list[h] = q for the flagged pairs (scan of the flags in `off`)
__global__ __launch_bounds__(256) void k_pair_hard_list(uint64_t np, const uint32_t* __restrict__ hard,
                                                        const uint32_t* __restrict__ off, uint32_t* __restrict__ list) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np && hard[q]) list[off[q]] = (uint32_t)q;
}

// One wave per pair of [p0, p0 + np) (grid-stride): the exact assignment of
// its cost matrix (copied to LDS when it fits, read from C otherwise).  LDS:
// potentials u, v, distances minv ((B + 1) doubles each), matching p,
// predecessors way, scanned flags ((B + 1) words each), the assignment and
// its inverse (B words each), then the matrix.
__global__ __launch_bounds__(SORT_BLOCK) void k_pair_lap(const double* __restrict__ Cg, uint32_t B, uint64_t p0,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ list_n, int cost_in_lds,
                                                         const uint32_t* __restrict__ pk,
                                                         const uint32_t* __restrict__ rev,
                                                         uint32_t* __restrict__ perm_index,
                                                         uint32_t* __restrict__ perms, uint32_t* __restrict__ bad) {
  extern __shared__ double lds[];
  const uint32_t lane = threadIdx.x;
  double* u = lds;
  double* v = u + (B + 1);
  double* minv = v + (B + 1);
  int32_t* p = reinterpret_cast<int32_t*>(minv + (B + 1));
  int32_t* way = p + (B + 1);
  int32_t* used = way + (B + 1);
  uint32_t* rowv = reinterpret_cast<uint32_t*>(used + (B + 1));
  uint32_t* colv = rowv + B;
  double* Cl = lds + 3 * (B + 1) + (3 * (B + 1) + 2 * B + 1) / 2;
  const double INF = __builtin_huge_val();
  const uint32_t BB = B * B;
  const uint64_t n_list = *list_n;
  for (uint64_t h = blockIdx.x; h < n_list; h += gridDim.x) {
    const uint64_t q0 = list[h];
    const double* C = Cg + q0 * BB;
    if (cost_in_lds) {
      for (uint32_t e = lane; e < BB; e += SORT_BLOCK) Cl[e] = C[e];
      C = Cl;
    }
    const uint32_t k = pk[p0 + q0];
    const uint32_t rv = rev[k];
    const bool has_rev = rv != 0xFFFFFFFFu;  // always, the lists being symmetric
    for (uint32_t q = lane; q <= B; q += SORT_BLOCK) {
      u[q] = 0;
      v[q] = 0;
      p[q] = 0;
      way[q] = 0;
    }
    __syncthreads();
    // successive shortest augmenting paths, rows 1..B; a cost matrix with
    // NaNs leaves no column to take, and the pair is abandoned
    bool stuck = false;
    for (uint32_t row = 1; row <= B && !stuck; ++row) {
      for (uint32_t q = lane; q <= B; q += SORT_BLOCK) {
        minv[q] = INF;
        used[q] = 0;
      }
      if (lane == 0) p[0] = (int32_t)row;
      __syncthreads();
      uint32_t j0 = 0;
      while (true) {
        if (lane == 0) used[j0] = 1;
        __syncthreads();
        const uint32_t i0 = (uint32_t)p[j0];
        double best = INF;
        uint32_t bj = 0xFFFFFFFFu;
        for (uint32_t q = 1 + lane; q <= B; q += SORT_BLOCK) {
          if (used[q]) continue;
          const double cur = C[(uint64_t)(i0 - 1) * B + (q - 1)] - u[i0] - v[q];
          if (cur < minv[q]) {
            minv[q] = cur;
            way[q] = (int32_t)j0;
          }
          if (minv[q] < best) {
            best = minv[q];
            bj = q;
          }
        }
        wave_argmin(best, bj);
        if (bj == 0xFFFFFFFFu) {  // wave-uniform
          stuck = true;
          break;
        }
        const double delta = best;
        __syncthreads();
        for (uint32_t q = lane; q <= B; q += SORT_BLOCK) {
          if (used[q]) {
            u[p[q]] += delta;
            v[q] -= delta;
          } else {
            minv[q] -= delta;
          }
        }
        __syncthreads();
        j0 = bj;
        if (p[j0] == 0) break;
      }
      if (lane == 0 && !stuck) {
        do {
          const uint32_t j1 = (uint32_t)way[j0];
          p[j0] = p[j1];
          j0 = j1;
        } while (j0);
      }
      __syncthreads();
    }
    if (stuck || !has_rev) {
      if (lane == 0) {
        atomicOr(bad, 1u);
        perm_index[k] = 0;
        if (has_rev) perm_index[rv] = 0;
      }
      __syncthreads();
      continue;
    }
    bool moved = false;
    for (uint32_t q = 1 + lane; q <= B; q += SORT_BLOCK) {
      rowv[p[q] - 1] = q - 1;
      colv[q - 1] = (uint32_t)(p[q] - 1);
      moved = moved || (uint32_t)(p[q] - 1) != q - 1;
    }
    moved = __any(moved);
    __syncthreads();
    if (!moved) {
      if (lane == 0) {
        perm_index[k] = 0;
        perm_index[rv] = 0;
      }
    } else {
      // fixed slots per pair (a counter shared by every pair serialised the
      // kernel on one address); the table is compacted afterwards
      const uint32_t slot = (uint32_t)(1 + 2 * (p0 + q0));
      for (uint32_t q = lane; q < B; q += SORT_BLOCK) {
        perms[(uint64_t)slot * B + q] = rowv[q];
        perms[(uint64_t)(slot + 1) * B + q] = colv[q];
      }
      if (lane == 0) {
        perm_index[k] = slot;
        perm_index[rv] = slot + 1;
      }
    }
    __syncthreads();  // the LDS is reused by the next pair
  }
}

// compaction of the table: moved[p] = 1 for pairs whose assignment is not
// the identity, then (after a scan) their two permutations move from slots
// 1 + 2p to 1 + 2 scan[p] and both keys are re-pointed
__global__ __launch_bounds__(256) void k_perm_moved(uint64_t np, const uint32_t* __restrict__ pk,
                                                    const uint32_t* __restrict__ perm_index,
                                                    uint32_t* __restrict__ moved) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) moved[p] = perm_index[pk[p]] != 0u ? 1u : 0u;
}
__global__ __launch_bounds__(256) void k_perm_compact(uint64_t np, uint32_t B, const uint32_t* __restrict__ pk,
                                                      const uint32_t* __restrict__ rev,
                                                      const uint32_t* __restrict__ scan,
                                                      uint32_t* __restrict__ perm_index,
                                                      const uint32_t* __restrict__ perms,
                                                      uint32_t* __restrict__ out) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const uint32_t k = pk[p];
  if (perm_index[k] == 0u) return;
  const uint64_t from = 1 + 2 * p, to = 1 + 2 * (uint64_t)scan[p];
  for (uint32_t q = 0; q < 2 * B; ++q) out[to * B + q] = perms[from * B + q];
  perm_index[k] = (uint32_t)to;
  perm_index[rev[k]] = (uint32_t)to + 1;
}

__global__ void k_identity_perm(uint32_t* perms, uint32_t B) {
  for (uint32_t q = threadIdx.x; q < B; q += blockDim.x) perms[q] = q;
}

struct DevBuf {
  std::vector<void*> held;
  ~DevBuf() {
    for (void* p : held) (void)hipFree(p);
  }
  template <typename T>
  T* get(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    held.push_back(p);
    return static_cast<T*>(p);
  }
  void release(void* p) {
    held.erase(std::remove(held.begin(), held.end(), p), held.end());
  }
};

}  // namespace
}  // namespace brille_amd

using namespace brille_amd;

extern "C" int brille_grid_sort(brille_grid* G, const brille_cost_info* values_cost,
                                const brille_cost_info* vectors_cost) {
  BRILLE_LOCK_GRID(G);
  if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
  if (!G->filled) return fail(BRILLE_ERR_NOT_FILLED, "The interpolation data must be filled before sorting.");
  DataConst& d = G->d;
  const brille_cost_info dflt{0, 0, {1.0, 1.0, 1.0}};
  const brille_cost_info& vc = values_cost ? *values_cost : dflt;
  const brille_cost_info& cc = vectors_cost ? *vectors_cost : dflt;
  if (d.B == 0 || d.V == 0) return BRILLE_OK;
  if (d.B > 4096) return fail(BRILLE_ERR_UNSUPPORTED, "sorting supports at most 4096 branches");
  if (d.V > 0xFFFFFFFFull) return fail(BRILLE_ERR_UNSUPPORTED, "sorting supports at most 2^32 - 1 vertices");
  const uint64_t V = d.V;
  const hipStream_t s = 0;
  DevBuf buf;
  const uint64_t units = 8 * G->n_cubes + 4 * G->n_tets;
  const uint64_t scan_words = (V + 1 + 4095) / 4096 + 1;
  uint32_t* cnt = buf.get<uint32_t>(V + 1);
  uint32_t* off = buf.get<uint32_t>(V + 1);
  uint32_t* part = buf.get<uint32_t>(scan_words);
  if (!cnt || !off || !part) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the sort work space");
  zero_words(cnt, (uint32_t)(V + 1), s);
  const unsigned ub = (unsigned)((units + 255) / 256), vb = (unsigned)((V + 255) / 256);
  if (units) hipLaunchKernelGGL(k_pair_emit<0>, dim3(ub), dim3(256), 0, s, G->g.cube_v, G->n_cubes, G->g.tet_v, G->n_tets, cnt, off, nullptr);
  if (hipMemcpyAsync(off, cnt, (V + 1) * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpyAsync failed");
  if (int rc = exclusive_scan_u32(off, V + 1, part, s); rc != BRILLE_OK) return rc;
  uint32_t total = 0;
  if (hipMemcpyAsync(&total, off + V, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "reading the pair count failed");
  uint32_t* adj = buf.get<uint32_t>(total);
  uint32_t* uniq = buf.get<uint32_t>(V + 1);
  uint32_t* koff = buf.get<uint32_t>(V + 1);
  if (!adj || !uniq || !koff) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the sort work space");
  zero_words(cnt, (uint32_t)(V + 1), s);
  if (units) hipLaunchKernelGGL(k_pair_emit<1>, dim3(ub), dim3(256), 0, s, G->g.cube_v, G->n_cubes, G->g.tet_v, G->n_tets, cnt, off, adj);
  zero_words(uniq, (uint32_t)(V + 1), s);
  hipLaunchKernelGGL(k_pair_unique, dim3(vb), dim3(256), 0, s, V, off, cnt, adj, uniq);
  if (hipMemcpyAsync(koff, uniq, (V + 1) * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpyAsync failed");
  if (int rc = exclusive_scan_u32(koff, V + 1, part, s); rc != BRILLE_OK) return rc;
  uint32_t nk32 = 0;
  if (hipMemcpyAsync(&nk32, koff + V, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "reading the key count failed");
  const uint64_t n_keys = nk32;
  // the new table: keys, per-key permutation index, permutations (at most
  // one per key plus the identity; trimmed below)
  uint64_t* keys = buf.get<uint64_t>(n_keys);
  uint32_t* nbr = buf.get<uint32_t>(n_keys);
  uint32_t* pidx = buf.get<uint32_t>(n_keys);
  uint32_t* perms = buf.get<uint32_t>((n_keys + 1) * (uint64_t)d.B);
  uint32_t* counter = buf.get<uint32_t>(2);
  if (!keys || !nbr || !pidx || !perms || !counter) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the permutation table");
  hipLaunchKernelGGL(k_pair_keys, dim3(vb), dim3(256), 0, s, V, off, adj, uniq, koff, keys, nbr);
  uint32_t* rev = buf.get<uint32_t>(n_keys);
  if (!rev) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the sort work space");
  if (n_keys)
    hipLaunchKernelGGL(k_pair_rev, dim3((unsigned)((n_keys + 255) / 256)), dim3(256), 0, s, V, n_keys, keys, nbr, koff,
                       rev);
  hipLaunchKernelGGL(k_identity_perm, dim3(1), dim3(256), 0, s, perms, d.B);
  zero_words(counter, 2, s);
  CostArgs ca{};
  ca.B = d.B;
  ca.sv = d.sv;
  ca.sc = d.sc;
  for (int q = 0; q < 3; ++q) {
    ca.ev[q] = d.ev[q];
    ca.ec[q] = d.ec[q];
    ca.vw[q] = vc.weights[q];
    ca.cw[q] = cc.weights[q];
  }
  ca.vs = vc.scalar_cost;
  ca.vv = vc.vector_cost;
  ca.cs = cc.scalar_cost;
  ca.cv = cc.vector_cost;
  ca.stride = d.rows64 ? d.stride64 : d.stride32;
  // the forward pairs (i < j) in key order
  uint32_t* fcnt = buf.get<uint32_t>(V + 1);
  if (!fcnt) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the sort work space");
  zero_words(fcnt, (uint32_t)(V + 1), s);
  hipLaunchKernelGGL(k_pair_fwd, dim3(vb), dim3(256), 0, s, V, koff, nbr, (const uint32_t*)nullptr, fcnt,
                     (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr);
  if (int rc = exclusive_scan_u32(fcnt, V + 1, part, s); rc != BRILLE_OK) return rc;
  uint32_t np32 = 0;
  if (hipMemcpyAsync(&np32, fcnt + V, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "reading the pair count failed");
  const uint64_t n_pairs = np32;
  uint32_t* pi = buf.get<uint32_t>(n_pairs);
  uint32_t* pj = buf.get<uint32_t>(n_pairs);
  uint32_t* pk = buf.get<uint32_t>(n_pairs);
  if (!pi || !pj || !pk) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the sort work space");
  hipLaunchKernelGGL(k_pair_fwd, dim3(vb), dim3(256), 0, s, V, koff, nbr, (const uint32_t*)fcnt, (uint32_t*)nullptr,
                     pi, pj, pk);
  // pairs in chunks whose cost matrices take at most 512 MB
  const uint64_t BB = (uint64_t)d.B * d.B;
  const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(n_pairs, (512ull << 20) / (BB * 8)));
  double* cost = buf.get<double>(std::max<uint64_t>(chunk, 1) * BB);
  uint32_t* hard = buf.get<uint32_t>(chunk + 1);
  uint32_t* hoff = buf.get<uint32_t>(chunk + 1);
  uint32_t* hpart = buf.get<uint32_t>((chunk + 1 + 4095) / 4096 + 1);
  uint32_t* hlist = buf.get<uint32_t>(chunk);
  if (!cost || !hard || !hoff || !hpart || !hlist) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the cost matrices");
  const size_t lds_state = 3 * (d.B + 1) * sizeof(double) + ((3 * (d.B + 1) + 2 * d.B + 1) / 2) * sizeof(double);
  const int cost_in_lds = lds_state + BB * sizeof(double) <= 48 * 1024 ? 1 : 0;
  const size_t lds = lds_state + (cost_in_lds ? BB * sizeof(double) : 0);
  const unsigned lap_blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(chunk, (uint64_t)std::max(G->n_cu, 1u) * 32));
  for (uint64_t p0 = 0; p0 < n_pairs; p0 += chunk) {
    const uint64_t np = std::min<uint64_t>(chunk, n_pairs - p0);
    const unsigned cb = (unsigned)((np * BB + 255) / 256);
    if (d.rows64)
      hipLaunchKernelGGL(k_pair_costs<double>, dim3(cb), dim3(256), 0, s, ca, d.rows64, pi, pj, p0, np, cost, counter + 1);
    else
      hipLaunchKernelGGL(k_pair_costs<float>, dim3(cb), dim3(256), 0, s, ca, d.rows32, pi, pj, p0, np, cost, counter + 1);
    const unsigned qb = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_pair_fast, dim3(qb), dim3(256), 0, s, (const double*)cost, d.B, p0, np, pk, rev, pidx, perms,
                       hard);
    // offsets of the flagged pairs (flags kept for the listing)
    if (hipMemcpyAsync(hoff, hard, np * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail(BRILLE_ERR_DEVICE, "hipMemcpyAsync failed");
    zero_words(hoff + np, 1, s);
    if (int rc = exclusive_scan_u32(hoff, np + 1, hpart, s); rc != BRILLE_OK) return rc;
    hipLaunchKernelGGL(k_pair_hard_list, dim3(qb), dim3(256), 0, s, np, hard, hoff, hlist);
    hipLaunchKernelGGL(k_pair_lap, dim3(lap_blocks), dim3(SORT_BLOCK), lds, s, (const double*)cost, d.B, p0,
                       (const uint32_t*)hlist, (const uint32_t*)(hoff + np), cost_in_lds, pk, rev, pidx, perms,
                       counter + 1);
  }
  // compact: the non-identity pairs' permutations, in pair order
  uint32_t* moved = buf.get<uint32_t>(n_pairs + 1);
  uint32_t* ppart = buf.get<uint32_t>((n_pairs + 1 + 4095) / 4096 + 1);
  if (!moved || !ppart) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the sort work space");
  zero_words(moved, (uint32_t)(n_pairs + 1), s);
  const unsigned pb = (unsigned)((n_pairs + 255) / 256);
  if (n_pairs) hipLaunchKernelGGL(k_perm_moved, dim3(pb), dim3(256), 0, s, n_pairs, pk, pidx, moved);
  if (int rc = exclusive_scan_u32(moved, n_pairs + 1, ppart, s); rc != BRILLE_OK) return rc;
  uint32_t n_moved = 0;
  if (hipMemcpyAsync(&n_moved, moved + n_pairs, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "sort failed");
  uint32_t* perms_out = buf.get<uint32_t>((1 + 2 * (uint64_t)n_moved) * d.B);
  if (!perms_out) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the permutation table");
  hipLaunchKernelGGL(k_identity_perm, dim3(1), dim3(256), 0, s, perms_out, d.B);
  if (n_pairs) hipLaunchKernelGGL(k_perm_compact, dim3(pb), dim3(256), 0, s, n_pairs, d.B, pk, rev, moved, pidx,
                                  perms, perms_out);
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "sort launch failed");
  uint32_t hc[2] = {0, 0};
  if (hipMemcpyAsync(hc, counter, 8, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "sort failed");
  if (hc[1]) return fail(BRILLE_ERR_NUMERIC, "Complex Hermitian angle cos(theta) > 1 while sorting");
  const uint32_t n_perms = 1 + 2 * n_moved;
  uint32_t* perms_fit = perms_out;
  // install: the previous sorted table goes (after the work still queued on
  // the device that may read it), a fill-given one stays in the payload
  // allocations until the next fill
  (void)hipDeviceSynchronize();
  for (void* p : G->sort_allocs) (void)hipFree(p);
  G->sort_allocs.clear();
  for (void* p : {(void*)keys, (void*)pidx, (void*)perms_fit}) {
    buf.release(p);
    G->sort_allocs.push_back(p);
  }
  d.n_keys = n_keys;
  d.keys = keys;
  d.perm_index = pidx;
  d.perms = perms_fit;
  G->n_perms = n_perms;
  // the bucketed row tables hold unpermuted rows; with a table the general
  // sums run and never read them
  d.cube_rows32 = nullptr;
  d.cube_rows64 = nullptr;
  d.wide_fac32 = nullptr;
  d.wide_fac64 = nullptr;
  ++G->fill_gen;
  return BRILLE_OK;
}

extern "C" int brille_grid_permutations(const brille_grid* G, uint64_t* n_keys, uint32_t* n_perms, uint64_t* keys,
                                        uint32_t* perm_index, uint32_t* perms) {
  BRILLE_LOCK_GRID(G);
  if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
  if (!G->filled) return fail(BRILLE_ERR_NOT_FILLED, "The interpolation data must be filled first.");
  const DataConst& d = G->d;
  if (n_keys) *n_keys = d.n_keys;
  if (n_perms) *n_perms = d.n_keys ? G->n_perms : 0;
  if (keys && d.n_keys && hipMemcpy(keys, d.keys, d.n_keys * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpy failed");
  if (perm_index && d.n_keys && hipMemcpy(perm_index, d.perm_index, d.n_keys * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpy failed");
  if (perms && d.n_keys && hipMemcpy(perms, d.perms, (uint64_t)G->n_perms * d.B * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpy failed");
  return BRILLE_OK;
}
