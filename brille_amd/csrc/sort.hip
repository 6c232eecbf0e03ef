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

// One wave per pair i < j of the key list, each block a contiguous run of
// keys (consecutive keys share their first vertex, whose rows then stay in
// LDS).  LDS: potentials u,
// v, distances minv ((B + 1) doubles each), matching p, predecessors way,
// scanned flags ((B + 1) words each), the assignment and its inverse (B words
// each), then (cost_in_lds) the B x B cost matrix; otherwise the matrix lives
// in cost_ws, B x B doubles per block.  The Dijkstra steps read a row of it
// each, so keeping it in LDS takes the L2 latency out of every step.
template <typename Real>
__global__ __launch_bounds__(SORT_BLOCK) void k_sort_pairs(CostArgs c, const Real* rows, uint64_t V, uint64_t n_keys,
                                                           const uint64_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ nbr,
                                                           const uint32_t* __restrict__ rev,
                                                           double* __restrict__ cost_ws, int cost_in_lds,
                                                           int rows_in_lds,
                                                           uint32_t* __restrict__ perm_index,
                                                           uint32_t* __restrict__ perms, uint32_t* __restrict__ counter,
                                                           uint32_t* __restrict__ bad) {
  extern __shared__ double lds[];
  const uint32_t B = c.B, lane = threadIdx.x;
  double* u = lds;
  double* v = u + (B + 1);
  double* minv = v + (B + 1);
  int32_t* p = reinterpret_cast<int32_t*>(minv + (B + 1));
  int32_t* way = p + (B + 1);
  int32_t* used = way + (B + 1);
  uint32_t* rowv = reinterpret_cast<uint32_t*>(used + (B + 1));
  uint32_t* colv = rowv + B;
  double* C = cost_in_lds ? lds + 3 * (B + 1) + (3 * (B + 1) + 2 * B + 1) / 2 : cost_ws + (uint64_t)blockIdx.x * B * B;
  // (rows_in_lds) the B rows of vertex i, then those of j, staged with
  // coalesced loads: the cost loops then read LDS instead of issuing a
  // dependent global load per element
  const uint32_t vrow = B * c.stride;  // reals per vertex
  Real* stage = reinterpret_cast<Real*>(C + (cost_in_lds ? (uint64_t)B * B : 0));
  const double INF = __builtin_huge_val();
  const uint64_t per = (n_keys + gridDim.x - 1) / gridDim.x;
  const uint64_t k_lo = (uint64_t)blockIdx.x * per, k_hi = k_lo + per < n_keys ? k_lo + per : n_keys;
  uint64_t staged_i = ~0ull;
  for (uint64_t k = k_lo; k < k_hi; ++k) {
    const uint64_t key = keys[k];
    const uint64_t i = key / V;
    const uint32_t j = nbr[k];
    if (i > j) continue;  // block-uniform
    const uint32_t rv = rev[k];
    const bool has_rev = rv != 0xFFFFFFFFu;  // always, the lists being symmetric
    const uint64_t r = rv;
    // costs
    const Real* ri = rows + i * vrow;
    const Real* rj = rows + (uint64_t)j * vrow;
    if (rows_in_lds) {
      if (i != staged_i)
        for (uint32_t q = lane; q < vrow; q += SORT_BLOCK) stage[q] = ri[q];
      for (uint32_t q = lane; q < vrow; q += SORT_BLOCK) stage[vrow + q] = rj[q];
      staged_i = i;
      __syncthreads();
      ri = stage;
      rj = stage + vrow;
    }
    bool ok = true;
    for (uint32_t e = lane; e < B * B; e += SORT_BLOCK) {
      const uint32_t a = e / B, b = e - a * B;
      C[e] = pair_cost(c, ri + a * c.stride, rj + b * c.stride, ok);
    }
    if (!ok) atomicOr(bad, 1u);
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
        if (has_rev) perm_index[r] = 0;
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
        perm_index[r] = 0;
      }
    } else {
      uint32_t slot = 0;
      if (lane == 0) slot = 1 + atomicAdd(counter, 2u);
      slot = __shfl(slot, 0);
      for (uint32_t q = lane; q < B; q += SORT_BLOCK) {
        perms[(uint64_t)slot * B + q] = rowv[q];
        perms[(uint64_t)(slot + 1) * B + q] = colv[q];
      }
      if (lane == 0) {
        perm_index[k] = slot;
        perm_index[r] = slot + 1;
      }
    }
    __syncthreads();  // the LDS and cost scratch are reused by the next pair
  }
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
  const unsigned blocks = (unsigned)std::min<uint64_t>(std::max<uint64_t>(n_keys, 1), (uint64_t)std::max(G->n_cu, 1u) * 32);
  // the LAP state, then the cost matrix when it fits in 48 KB of LDS
  const size_t lds_state = 3 * (d.B + 1) * sizeof(double) + ((3 * (d.B + 1) + 2 * d.B + 1) / 2) * sizeof(double);
  const size_t lds_cost = (size_t)d.B * d.B * sizeof(double);
  const int cost_in_lds = lds_state + lds_cost <= 48 * 1024 ? 1 : 0;
  double* cost_ws = cost_in_lds ? nullptr : buf.get<double>((uint64_t)blocks * d.B * d.B);
  if (!cost_in_lds && !cost_ws) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the cost matrices");
  const size_t rs = d.rows64 ? sizeof(double) : sizeof(float);
  const size_t lds_rows = 2 * (size_t)d.B * (d.rows64 ? d.stride64 : d.stride32) * rs;
  const int rows_in_lds = cost_in_lds && lds_state + lds_cost + lds_rows <= 60 * 1024 ? 1 : 0;
  const size_t lds = lds_state + (cost_in_lds ? lds_cost : 0) + (rows_in_lds ? lds_rows : 0);
  if (n_keys) {
    if (d.rows64) {
      ca.stride = d.stride64;
      hipLaunchKernelGGL(k_sort_pairs<double>, dim3(blocks), dim3(SORT_BLOCK), lds, s, ca, d.rows64, V, n_keys, keys, nbr,
                         rev, cost_ws, cost_in_lds, rows_in_lds, pidx, perms, counter, counter + 1);
    } else {
      ca.stride = d.stride32;
      hipLaunchKernelGGL(k_sort_pairs<float>, dim3(blocks), dim3(SORT_BLOCK), lds, s, ca, d.rows32, V, n_keys, keys, nbr,
                         rev, cost_ws, cost_in_lds, rows_in_lds, pidx, perms, counter, counter + 1);
    }
  }
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "sort launch failed");
  uint32_t hc[2] = {0, 0};
  if (hipMemcpyAsync(hc, counter, 8, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "sort failed");
  if (hc[1]) return fail(BRILLE_ERR_NUMERIC, "Complex Hermitian angle cos(theta) > 1 while sorting");
  const uint32_t n_perms = 1 + hc[0];
  uint32_t* perms_fit = buf.get<uint32_t>((uint64_t)n_perms * d.B);
  if (!perms_fit) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the permutation table");
  if (hipMemcpy(perms_fit, perms, (uint64_t)n_perms * d.B * 4, hipMemcpyDeviceToDevice) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpy failed");
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
