// Binned ir_interpolate_at for large calls.
//
// With Q spread over the whole zone, the eight payload rows a point gathers
// are unrelated to those of its neighbours in the input, and on a payload
// larger than the L2 nearly every row comes from HBM.  This pipeline first
// orders the points by trellis node range, so that the points one tile sums
// share their vertices and the rows hit in cache:
//   k_bin_fold     stage (a) + node subscript, one lane per point; a 32-byte
//                  FoldRec per point in input order and a per-block histogram
//                  of node-range buckets (failed points get the last bucket);
//   k_scan_*       exclusive scan of the block histograms (bucket-major);
//   k_bin_scatter  each block moves its FoldRecs to its reserved slots of each
//                  bucket (a SortRec adds the input index);
//   k_bin_sums     the rest of stage (b) and stages (c1, c2) on bucket-ordered
//                  tiles, writing output rows at the input index.
// Within a bucket the order is whatever the LDS atomics give; every point is
// computed independently, so the outputs do not depend on it.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "grid_state.hpp"

namespace brille_amd {

constexpr int BIN_BLOCK = 256;
constexpr int BIN_PPT = 8;                            // points per thread in the bucketing kernels
constexpr uint32_t BIN_SPAN = BIN_BLOCK * BIN_PPT;   // points per bucketing block
constexpr uint64_t BIN_CHUNK = 1ull << 24;           // points per pass of the pipeline
constexpr uint32_t SCAN_SEG = 4096;                   // elements per scan block (256 x 16)

// point-group operations the binned sums keep a Gamma frame for
constexpr uint32_t kMaxOps = 48;

struct FoldRec {
  double q[3];
  uint32_t node;
  uint32_t meta;  // rot | invrot << 8 | status << 16 | kMetaPoly
};
// the point is on the proposal list: it folded into a polyhedron node of at
// least kPropMinCand tetrahedra (in smaller nodes the exact scan costs no
// more than verifying a proposal, and the network is not run)
constexpr uint32_t kMetaPoly = 1u << 24;
// the same flag on the node words the node-order pass reads (node indices
// stay below 2^31, which the launch checks) and on the perm entries it
// writes (positions within a pass stay below 2^31)
constexpr uint32_t kNodePoly = 1u << 31;
constexpr uint32_t kPermPoly = 1u << 31;
constexpr uint32_t kPropMinCand = 8;
__device__ __forceinline__ uint32_t meta_status(uint32_t meta) { return (meta >> 16) & 255u; }
struct SortRec {
  double q[3];
  uint32_t node;
  uint32_t meta;
  uint32_t index;  // input index within the pass
  uint32_t hint;   // proposed tetrahedron (kNoTet: none)
};

__device__ __forceinline__ uint32_t bucket_of(uint32_t node, uint32_t meta, uint32_t nb, uint64_t n_nodes) {
  return meta_status(meta) != BRILLE_PT_OK ? nb : (uint32_t)(((uint64_t)node * nb) / n_nodes);
}

// mark_poly: flag (kMetaPoly) the points for the learned proposal, those
// that folded into polyhedron nodes of at least kPropMinCand tetrahedra; the
// node-order pass lists them
// Six waves per SIMD (a few spilled registers) hide more of the FP64 latency
// than the four that 124 VGPRs allow: 9 % less time.
template <bool Fast>
__global__ __launch_bounds__(BIN_BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_bin_fold(GridConst g, const double* Q, uint32_t n, int fold,
                                                        FoldRec* __restrict__ rec, uint32_t* __restrict__ hist,
                                                        uint32_t nblk, uint32_t nb,
                                                        ErrRec* __restrict__ err, uint64_t base, int mark_poly) {
  extern __shared__ uint32_t h[];  // nb + 1 counters
  for (uint32_t k = threadIdx.x; k <= nb; k += BIN_BLOCK) h[k] = 0;
  __syncthreads();
  const uint32_t b0 = blockIdx.x * BIN_SPAN;
  for (int i = 0; i < BIN_PPT; ++i) {
    const uint32_t p = b0 + i * BIN_BLOCK + threadIdx.x;
    if (p < n) {
      double Qp[3] = {Q[3 * (uint64_t)p], Q[3 * (uint64_t)p + 1], Q[3 * (uint64_t)p + 2]};
      Located L;
      fold_node<Fast>(g, Qp, fold, L);
      if (L.status != BRILLE_PT_OK) report_point(err, base + p, L.status);
      bool poly = false;
      if (mark_poly && L.status == BRILLE_PT_OK && g.node_type[L.node] == BRILLE_NODE_POLY) {
        const uint32_t pi = g.node_index[L.node];
        poly = g.poly_off[pi + 1] - g.poly_off[pi] >= kPropMinCand;
      }
      FoldRec r;
      r.q[0] = L.q[0]; r.q[1] = L.q[1]; r.q[2] = L.q[2];
      r.node = L.node;
      r.meta = (L.rot & 255u) | ((L.invrot & 255u) << 8) | ((uint32_t)L.status << 16) | (poly ? kMetaPoly : 0u);
      rec[p] = r;
      atomicAdd(&h[bucket_of(r.node, r.meta, nb, g.n_nodes)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k <= nb; k += BIN_BLOCK) hist[(uint64_t)k * nblk + blockIdx.x] = h[k];
}

// exclusive scan of 256 values held one per thread; returns the thread's
// exclusive prefix, *total gets the sum
template <int NT = BIN_BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[NT / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    if ((uint32_t)w < wave) before += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__global__ __launch_bounds__(BIN_BLOCK) void k_scan_reduce(const uint32_t* a, uint64_t m, uint32_t* part) {
  const uint64_t s0 = (uint64_t)blockIdx.x * SCAN_SEG;
  uint32_t v = 0;
  for (uint32_t j = 0; j < SCAN_SEG / BIN_BLOCK; ++j) {
    const uint64_t i = s0 + (uint64_t)threadIdx.x * (SCAN_SEG / BIN_BLOCK) + j;
    if (i < m) v += a[i];
  }
  uint32_t total;
  (void)block_exclusive_scan(v, &total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// one block: exclusive scan of the segment sums in place
__global__ __launch_bounds__(BIN_BLOCK) void k_scan_top(uint32_t* part, uint32_t np) {
  const uint32_t per = (np + BIN_BLOCK - 1) / BIN_BLOCK;
  const uint32_t i0 = threadIdx.x * per;
  uint32_t v = 0;
  for (uint32_t j = 0; j < per; ++j)
    if (i0 + j < np) v += part[i0 + j];
  uint32_t total;
  uint32_t run = block_exclusive_scan(v, &total);
  for (uint32_t j = 0; j < per; ++j)
    if (i0 + j < np) {
      const uint32_t t = part[i0 + j];
      part[i0 + j] = run;
      run += t;
    }
}

__global__ __launch_bounds__(BIN_BLOCK) void k_scan_apply(uint32_t* a, uint64_t m, const uint32_t* part) {
  constexpr uint32_t PER = SCAN_SEG / BIN_BLOCK;
  const uint64_t i0 = (uint64_t)blockIdx.x * SCAN_SEG + (uint64_t)threadIdx.x * PER;
  uint32_t v[PER];
  uint32_t s = 0;
#pragma unroll
  for (uint32_t j = 0; j < PER; ++j) {
    v[j] = i0 + j < m ? a[i0 + j] : 0u;
    s += v[j];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, &total) + part[blockIdx.x];
#pragma unroll
  for (uint32_t j = 0; j < PER; ++j) {
    if (i0 + j < m) a[i0 + j] = run;
    run += v[j];
  }
}

__global__ __launch_bounds__(BIN_BLOCK) void k_bin_scatter(const FoldRec* rec, uint32_t n, const uint32_t* off,
                                                           uint32_t nblk, uint32_t nb, uint64_t n_nodes, SortRec* out,
                                                           uint32_t* out_node) {
  extern __shared__ uint32_t s[];  // base[nb + 1], count[nb + 1]
  uint32_t* base = s;
  uint32_t* cnt = s + nb + 1;
  for (uint32_t k = threadIdx.x; k <= nb; k += BIN_BLOCK) {
    base[k] = off[(uint64_t)k * nblk + blockIdx.x];
    cnt[k] = 0;
  }
  __syncthreads();
  const uint32_t b0 = blockIdx.x * BIN_SPAN;
  constexpr int GRP = 8;  // records requested before any is placed
  for (int i0 = 0; i0 < BIN_PPT; i0 += GRP) {
    FoldRec rr[GRP];
#pragma unroll
    for (int i = 0; i < GRP; ++i) {
      const uint32_t p = b0 + (i0 + i) * BIN_BLOCK + threadIdx.x;
      if (p < n) rr[i] = rec[p];
    }
#pragma unroll
    for (int i = 0; i < GRP; ++i) {
    const uint32_t p = b0 + (i0 + i) * BIN_BLOCK + threadIdx.x;
    if (p < n) {
      const FoldRec& r = rr[i];
      const uint32_t k = bucket_of(r.node, r.meta, nb, n_nodes);
      const uint32_t j = atomicAdd(&cnt[k], 1u);
      SortRec o;
      o.q[0] = r.q[0]; o.q[1] = r.q[1]; o.q[2] = r.q[2];
      o.node = r.node;
      o.meta = r.meta;
      o.index = p;
      o.hint = kNoTet;
      out[base[k] + j] = o;
      out_node[base[k] + j] = r.node | ((r.meta & kMetaPoly) ? kNodePoly : 0u);
    }
    }
  }
}

// Node order inside each bucket: one block per bucket counts its points per
// trellis node in LDS and writes perm, the positions of the bucket's records
// in node order, so that the points of one cube are summed together and its
// rows are read from L1.  Within a node the order is whatever the LDS atomics
// give (the points are independent).  Buckets spanning more than kPermNodes
// nodes, and the failed-point bucket, keep their order.  Each entry carries
// the record's proposal flag (kPermPoly, from kNodePoly).
constexpr uint32_t kPermNodes = 4096;
constexpr int PERM_BLOCK = 1024;
__global__ __launch_bounds__(PERM_BLOCK) void k_bucket_perm(const uint32_t* node, uint32_t m, const uint32_t* off,
                                                           uint32_t nblk, uint32_t nb, uint64_t n_nodes,
                                                           uint32_t* perm) {
  __shared__ uint32_t cnt[kPermNodes];
  const uint32_t b = blockIdx.x;
  const uint32_t s0 = off[(uint64_t)b * nblk];
  const uint32_t s1 = b + 1 <= nb ? off[(uint64_t)(b + 1) * nblk] : m;
  const uint64_t lo = ((uint64_t)b * n_nodes + nb - 1) / nb;
  const uint64_t hi = b < nb ? ((uint64_t)(b + 1) * n_nodes + nb - 1) / nb : lo;
  const uint32_t nn = (uint32_t)(hi - lo);
  if (b == nb || nn == 0 || nn > kPermNodes) {
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += PERM_BLOCK) perm[i] = i | (node[i] & kNodePoly);
    return;
  }
  for (uint32_t k = threadIdx.x; k < nn; k += PERM_BLOCK) cnt[k] = 0;
  __syncthreads();
  for (uint32_t i = s0 + threadIdx.x; i < s1; i += PERM_BLOCK) atomicAdd(&cnt[(node[i] & ~kNodePoly) - (uint32_t)lo], 1u);
  __syncthreads();
  // exclusive scan of cnt[0, nn): each thread a contiguous run, then the runs
  const uint32_t per = (nn + PERM_BLOCK - 1) / PERM_BLOCK;
  const uint32_t k0 = threadIdx.x * per, k1 = k0 + per < nn ? k0 + per : nn;
  uint32_t run = 0;
  for (uint32_t k = k0; k < k1; ++k) run += cnt[k];
  uint32_t total;
  uint32_t acc = s0 + block_exclusive_scan<PERM_BLOCK>(run, &total);
  for (uint32_t k = k0; k < k1; ++k) {
    const uint32_t c = cnt[k];
    cnt[k] = acc;
    acc += c;
  }
  __syncthreads();
  for (uint32_t i = s0 + threadIdx.x; i < s1; i += PERM_BLOCK) {
    const uint32_t nd = node[i];
    perm[atomicAdd(&cnt[(nd & ~kNodePoly) - (uint32_t)lo], 1u)] = i | (nd & kNodePoly);
  }
}

// The flagged records (kPermPoly) of perm, listed in perm order: each block
// compacts its 4096 consecutive entries in LDS and appends them with one
// atomic, so the list runs in node order within each block's share.
constexpr int LIST_PPT = 16;
constexpr uint32_t LIST_SPAN = BIN_BLOCK * LIST_PPT;
__global__ __launch_bounds__(BIN_BLOCK) void k_poly_list(const uint32_t* perm, uint32_t m, uint32_t* __restrict__ list,
                                                        uint32_t* __restrict__ count) {
  __shared__ uint32_t s_list[LIST_SPAN];
  __shared__ uint32_t s_n, s_at;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t b0 = (uint64_t)blockIdx.x * LIST_SPAN;
  // each wave a contiguous quarter of the span, 64 entries at a time, so the
  // LDS list keeps perm order within each wave's quarter
  __shared__ uint32_t s_wn[BIN_BLOCK / 64];
  uint32_t local = 0;
  // pass 1: count per wave
  const uint64_t w0 = b0 + (uint64_t)wave * (LIST_SPAN / (BIN_BLOCK / 64));
  for (uint32_t k = 0; k < LIST_SPAN / BIN_BLOCK; ++k) {
    const uint64_t j = w0 + (uint64_t)k * 64u + lane;
    const bool poly = j < m && (perm[j] & kPermPoly) != 0u;
    local += (uint32_t)__popcll(__ballot(poly));
  }
  if (lane == 0) s_wn[wave] = local;
  __syncthreads();
  uint32_t at = 0;
  for (uint32_t w = 0; w < wave; ++w) at += s_wn[w];
  // pass 2: place, in order within the wave's quarter
  for (uint32_t k = 0; k < LIST_SPAN / BIN_BLOCK; ++k) {
    const uint64_t j = w0 + (uint64_t)k * 64u + lane;
    const uint32_t e = j < m ? perm[j] : 0u;
    const bool poly = (e & kPermPoly) != 0u;
    const unsigned long long mask = __ballot(poly);
    if (poly) s_list[at + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = e & ~kPermPoly;
    at += (uint32_t)__popcll(mask);
  }
  if (threadIdx.x == BIN_BLOCK - 1) s_n = at;  // the last wave ends at the block's total
  __syncthreads();
  if (threadIdx.x == 0) s_at = s_n ? atomicAdd(count, s_n) : 0u;
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < s_n; k += BIN_BLOCK) list[s_at + k] = s_list[k];
}

// The learned proposal for the listed records: each wave takes 16 list
// entries at a time (mostly of one node, so the candidate rows the chunk
// reads are mostly the same rows), stages their features and candidate
// ranges in its LDS slice, runs the network and writes the proposed
// tetrahedron to the record's hint.  Grid-stride over the list, whose length
// is read on the device.
__global__ __launch_bounds__(256) void k_bin_propose(GridConst g, ProposalConst pc, SortRec* srt,
                                                     const uint32_t* list, const uint32_t* count) {
  // chunks of 16 listed points per wave, four waves per block, grid-stride
  // over the list; blocks past the list's end leave before copying weights
  const uint32_t n = *count;
  if (blockIdx.x * 64u >= n) return;
  __shared__ float s_feat[4][48];
  __shared__ uint32_t s_t0[4][16], s_nc[4][16], s_out[4][16];
  // the first two layers' weights, from LDS copies
  __shared__ float s_w1[3 * kPropH], s_b1[kPropH], s_b2[kPropH];
  __shared__ __attribute__((aligned(16))) uint16_t s_w2[kPropH * kPropH];
  for (uint32_t i = threadIdx.x; i < 3 * kPropH; i += 256) s_w1[i] = pc.W1[i];
  for (uint32_t i = threadIdx.x; i < kPropH; i += 256) {
    s_b1[i] = pc.b1[i];
    s_b2[i] = pc.b2[i];
  }
  for (uint32_t i = threadIdx.x; i < kPropH * kPropH / 8; i += 256)
    reinterpret_cast<uint4*>(s_w2)[i] = reinterpret_cast<const uint4*>(pc.W2)[i];
  __syncthreads();
  const PropWeights<LdsF, LdsH> W{(LdsF)s_w1, (LdsF)s_b1, (LdsF)s_b2, (LdsH)s_w2};
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  // wave-uniform loop; no block-wide barrier follows
  for (uint32_t c0 = (blockIdx.x * 4u + wave) * 16u; c0 < n; c0 += gridDim.x * 64u) {
  const uint32_t k = n - c0 < 16u ? n - c0 : 16u;
  uint32_t rec = 0;
  if (lane < k) {
    rec = list[c0 + lane];
    const SortRec& r = srt[rec];
    const double q[3] = {r.q[0], r.q[1], r.q[2]};
    const uint32_t nd = r.node;
    double x[3];
    mat_vec(x, g.outer_to_xyz, q);
    float f[3];
    prop_features(g, x, nd, f);
    s_feat[wave][3 * lane] = f[0];
    s_feat[wave][3 * lane + 1] = f[1];
    s_feat[wave][3 * lane + 2] = f[2];
    const uint32_t pi = g.node_index[nd];
    const uint32_t a = g.poly_off[pi];
    s_t0[wave][lane] = a;
    s_nc[wave][lane] = g.poly_off[pi + 1] - a;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  propose16_inline(pc, W, s_feat[wave], s_t0[wave], s_nc[wave], k, s_out[wave]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < k) srt[rec].hint = s_out[wave][lane];
  // the next chunk overwrites this wave's LDS slices
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// XCD-aware tile order: the hardware hands consecutive blocks to the eight
// XCDs in turn, each with its own L2.  Block b runs on XCD b % 8, so giving XCD
// x the contiguous tiles [start(x), start(x) + count(x)) keeps a bucket's
// records and rows in one L2 instead of fetching them into all eight.  The
// map is a bijection for any grid size.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
  const uint32_t q = nb / kXcds, r = nb % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

// (The FP32 wide-row sums run at two waves per SIMD, over 200 VGPRs: held
// to three waves (168 VGPRs, some spilled) the wave-per-unit walk gained
// 12 %, the atom-major walk lost 8 %; four waves spill too much.)
template <typename Real, int L, int SC, int SV, bool Prop>
__global__ __launch_bounds__(TILE) void k_bin_sums(GridConst g, DataConst d, ProposalConst pc, const SortRec* srt,
                                                    const uint32_t* perm, uint32_t n, void* __restrict__ vals,
                                                    void* __restrict__ vecs, brille_aux_out aux,
                                                    ErrRec* __restrict__ err, uint64_t base, uint32_t* rest,
                                                    uint32_t* rest_n) {
  using C = typename Cplx<Real>::T;
  using S = TileShape<L, SC, TILE, Real>;
  __shared__ PointRec<Real> s_rec[TILE];
  // wide rows: the phase table of the node-ordered Gamma sums instead
  constexpr bool kWide = L == LAYOUT_WIDE;
  constexpr uint32_t kPh =
      kWide ? (SC == kWideAtom ? kWideAtomPhaseBytes : kWidePhaseBytes) / sizeof(C) : TILE * S::NPH;
  __shared__ C s_ph[kPh];
  __shared__ Real s_wrot[kWide ? 9 * kMaxOps : 1];
  __shared__ uint8_t s_wdst[kWide ? kWideAtoms * kMaxOps : 1];
  // Gamma layout: one frame per point-group operation instead of one per
  // point (12 KB less LDS in FP32, 20 KB in FP64)
  constexpr bool kOpFrames = L == LAYOUT_GAMMA;
  __shared__ double s_q[kOpFrames ? 1 : S::NQ];
  __shared__ GammaFrame<Real> s_frames[kOpFrames ? kMaxOps : 1];
  if (kOpFrames && threadIdx.x < g.G) fill_frame<Real, SC>(g, d, threadIdx.x, s_frames[threadIdx.x]);
  __shared__ uint32_t s_out[TILE];
  __shared__ C s_wide[S::NW];
  __shared__ uint32_t s_tried, s_verified;
  if (Prop && threadIdx.x == 0) {
    s_tried = 0;
    s_verified = 0;
  }
  // wide rows, kWideAtom: the located points without a table block are
  // listed in rest[0 .. *rest_n) (record indices) for the kWideRest launch,
  // which takes its points from that list
  constexpr bool kAtom = L == LAYOUT_WIDE && SC == kWideAtom;
  constexpr bool kRest = L == LAYOUT_WIDE && SC == kWideRest;
  if constexpr (kRest) n = min(n, *rest_n);
  // kWideRest: a few blocks (the launch has no count of the list) walk the
  // listed points tile by tile; every other launch has one tile per block
  const uint32_t first_tile = kRest ? blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  for (uint32_t tile = first_tile; (uint64_t)tile * TILE < n; tile += gridDim.x) {
    if (tile != first_tile) __syncthreads();  // the previous tile's records are read
    const uint32_t tile0 = tile * TILE;
    const uint32_t tid = threadIdx.x;
    const uint32_t p = tile0 + tid;
    // wide rows: blockIdx.y picks a group of modes (gridDim.y groups; the
    // first one reports the points' errors and aux outputs)
    const bool lead = blockIdx.y == 0;
    const uint32_t per_group = (d.B + gridDim.y - 1) / gridDim.y;
    const uint32_t b_lo = blockIdx.y * per_group, b_hi = b_lo + per_group;
    const bool want_phase = d.vec_rot == BRILLE_ROT_GAMMA && d.ec[1] > 0;
    const uint32_t n_atoms = d.ec[1] / 3u;
    const bool ph_in_lds = L == LAYOUT_GAMMA || (want_phase && n_atoms <= 4);
    WideGammaLds<Real> wg;
    if (kWide && want_phase) {
      wg.ph = s_ph;
      wg.rot = s_wrot;
      wg.dst = s_wdst;
      wg.cap = kPh;
      wide_gamma_ops<Real, TILE>(g, d, wg);
    }
    __shared__ uint32_t s_wave_tab[TILE / 64];
    if (Prop) __syncthreads();
    const bool live = p < n;
    SortRec sr;
    Located P;
    P.status = BRILLE_PT_OK;
    P.cube = kNoCube;
    uint32_t ridx = 0;
    if (live) {
      ridx = kRest ? rest[p] : (perm[p] & ~kPermPoly);
      sr = srt[ridx];
      P.q[0] = sr.q[0]; P.q[1] = sr.q[1]; P.q[2] = sr.q[2];
      P.tau[0] = P.tau[1] = P.tau[2] = 0;
      P.node = sr.node;
      P.rot = sr.meta & 255u;
      P.invrot = (sr.meta >> 8) & 255u;
      P.status = (int)meta_status(sr.meta);
      P.tet = kNoTet;
      P.nslot = 0;
      P.nkept = 0;
      if (P.status == BRILLE_PT_OK) {
        if constexpr (Prop) {
          bool verified = false;
          node_slots(g, P, sr.hint, pc.margin, &verified);
          if (sr.hint != kNoTet) {
            atomicAdd(&s_tried, 1u);
            if (verified) atomicAdd(&s_verified, 1u);
          }
        } else {
          node_slots(g, P);
        }
        // failed fold-ins were reported by k_bin_fold
        if (P.status != BRILLE_PT_OK && lead) report_point(err, base + sr.index, P.status);
        if (kAtom && lead && P.status == BRILLE_PT_OK && table_row(d, P) == kNoCube) rest[atomicAdd(rest_n, 1u)] = ridx;
      }
    }
    // Stable partition of the tile: points the cube factor table serves first,
    // so that few waves of the sums mix the two kinds of point.
    const bool table = (L == LAYOUT_GAMMA || L == LAYOUT_SCALAR) && cube_rows<Real>(d) != nullptr;
    const bool tab = table && live && P.status == BRILLE_PT_OK && table_row(d, P) != kNoCube;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const unsigned long long bal = __ballot(tab);
    const uint32_t rank = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave_tab[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0, n_tab = 0;
  #pragma unroll
    for (uint32_t w = 0; w < TILE / 64; ++w) {
      const uint32_t c = s_wave_tab[w];
      before += w < wave ? c : 0u;
      n_tab += c;
    }
    const uint32_t slot = tab ? before + rank : n_tab + (tid - before - rank);
    {
      PointRec<Real>& r = s_rec[slot];
      r.nslot = 0;
      if (live) {
        s_out[slot] = sr.index;
        if (P.status == BRILLE_PT_OK)
          store_record<Real, L, SC>(g, d, P, r, s_ph, s_q, slot, ph_in_lds, n_atoms, !kOpFrames);
        if (lead && !kRest) write_aux(aux, sr.index, P);
      }
    }
    __syncthreads();
    if (Prop && lead && threadIdx.x == 0 && s_tried) {
      atomicAdd(&pc.stats[0], (unsigned long long)s_tried);
      atomicAdd(&pc.stats[1], (unsigned long long)s_verified);
    }
    const uint32_t npts = n - tile0 < (uint32_t)TILE ? n - tile0 : (uint32_t)TILE;
    // the one-lane-per-unit sums see only the points the table does not serve:
    // fewer rows in flight keep the registers for occupancy
    constexpr int GRP = sizeof(Real) == 4 ? 2 : 1;
    tile_sums<Real, true, L, SC, SV, TILE, GRP>(g, d, s_rec, s_ph, s_q, s_out, 0, npts, ph_in_lds, vals, vecs, s_wide,
                                                table, n_tab, kOpFrames ? s_frames : nullptr, true, wg, b_lo, b_hi);

    if (!kRest) break;
  }
}

// The row table: thread t = (block row, mode b) phases the row of a cube's or
// tetrahedron's slot-x vertex against the row of its slot-0 vertex with the
// sums' own functions, so a point summed from the table gets the same bits.
template <typename Real, int SC>
__global__ __launch_bounds__(256) void k_table_rows(GridConst g, DataConst d, uint64_t n_cube_units,
                                                    uint64_t n_units, Real* out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_units) return;
  uint32_t pv, sv;
  uint32_t b;
  if (t < n_cube_units) {
    const uint64_t cx = t / d.B;
    b = (uint32_t)(t - cx * d.B);
    const uint32_t* v = g.cube_v + 8 * (cx >> 3);
    pv = v[7];
    sv = v[7 - (int)(cx & 7u)];
  } else {
    const uint64_t tx = (t - n_cube_units) / d.B;
    b = (uint32_t)(t - n_cube_units - tx * d.B);
    const uint32_t* v = g.tet_v + 4 * (tx >> 2);
    pv = v[0];
    sv = v[tx & 3u];
  }
  const Real* base = payload_rows<Real>(d);
  const uint64_t stride = payload_stride<Real>(d);
  Real pr[2 * SC];
  load_row<Real, 2 * SC>(pr, base + ((uint64_t)pv * d.B + b) * stride);
  const Real* src = base + ((uint64_t)sv * d.B + b) * stride;
  Real row[2 * SC];
  load_row<Real, 2 * SC>(row, src);
  const typename Cplx<Real>::T e = slot_antiphase<Real, SC>(pr, row);
  Real* dst = out + t * stride;
  for (int k = 0; k < SC; ++k) phase_elem<Real>(e, row[2 * k], row[2 * k + 1], &dst[2 * k], &dst[2 * k + 1]);
  for (uint32_t k = 2 * SC; k < stride; ++k) dst[k] = src[k];  // values and zero padding
}

// The wide factor table: wave w = (block row, mode b) computes the antiphases
// of a cube's eight or a tetrahedron's four slot rows against its slot-0 row
// with the sums' own wave-wide function (same lanes, same reductions), so a
// point summed from the table gets the same bits.
template <typename Real>
__global__ __launch_bounds__(256) void k_wide_factors(GridConst g, DataConst d, uint64_t n_cube_units,
                                                      uint64_t n_units, typename Cplx<Real>::T* out) {
  using C = typename Cplx<Real>::T;
  const uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n_units) return;  // wave-uniform
  const uint32_t l = threadIdx.x & 63u;
  const uint32_t sc = d.sc;
  const bool h0 = l < sc, h1 = l + 64u < sc;
  uint32_t v[8], ns, b;
  uint64_t trow;
  if (t < n_cube_units) {
    const uint64_t c = t / d.B;
    b = (uint32_t)(t - c * d.B);
#pragma unroll
    for (int x = 0; x < 8; ++x) v[x] = g.cube_v[8 * c + 7 - x];
    ns = 8;
    trow = c * 8 * d.B;
  } else {
    const uint64_t tt = (t - n_cube_units) / d.B;
    b = (uint32_t)(t - n_cube_units - tt * d.B);
#pragma unroll
    for (int x = 0; x < 8; ++x) v[x] = g.tet_v[4 * tt + (x < 4 ? x : 0)];
    ns = 4;
    trow = d.tab_tet_row + tt * 4 * d.B;
  }
  const C zero = Elem<C>::zero();
  C v0[8], v1[8], e[8];
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const C* rc = reinterpret_cast<const C*>(row_at<Real>(d, v[x], b));
    v0[x] = h0 ? rc[l] : zero;
    v1[x] = h1 ? rc[l + 64] : zero;
  }
  wide_antiphases<Real, 8>(v0[0], v1[0], v0, v1, e, sc);
  if (l == 0) {
#pragma unroll
    for (int x = 0; x < 8; ++x)
      if ((uint32_t)x < ns) out[trow + (uint64_t)x * d.B + b] = e[x];
  }
}

using SumsKernel = void (*)(GridConst, DataConst, ProposalConst, const SortRec*, const uint32_t*, uint32_t, void*,
                           void*, brille_aux_out, ErrRec*, uint64_t, uint32_t*, uint32_t*);

// Gamma rows too long for one lane that the atom-major walk takes whole: the
// derived antiphase factor table exists and there is at most one value per atom
static bool wide_by_atom(const DataConst& d, int precision) {
  const bool fac = precision == BRILLE_FP32 ? d.wide_fac32 != nullptr : d.wide_fac64 != nullptr;
  return payload_layout(d) == LAYOUT_GAMMA && d.vec_rot == BRILLE_ROT_GAMMA && d.ec[1] > 0 && fac &&
         d.sv <= d.sc / 3u;
}

template <typename Real>
static SumsKernel pick_sums(const DataConst& d, bool prop) {
  const int L = payload_layout(d);
  if (prop) {
    // the learned proposal is instantiated for the two-atom Gamma payload and the general layout
    if (L == LAYOUT_GAMMA && d.sv == 1 && d.sc == 6) return k_bin_sums<Real, LAYOUT_GAMMA, 6, 1, true>;
    return k_bin_sums<Real, LAYOUT_GENERIC, 0, 0, true>;
  }
  if (L == LAYOUT_GAMMA && d.sv == 1) {
    switch (d.sc) {
      case 3: return k_bin_sums<Real, LAYOUT_GAMMA, 3, 1, false>;
      case 6: return k_bin_sums<Real, LAYOUT_GAMMA, 6, 1, false>;
      case 9: return k_bin_sums<Real, LAYOUT_GAMMA, 9, 1, false>;
      case 12: return k_bin_sums<Real, LAYOUT_GAMMA, 12, 1, false>;
      default: break;
    }
  }
  if (L == LAYOUT_SCALAR && d.sv == 1) {
    switch (d.sc) {
      case 1: return k_bin_sums<Real, LAYOUT_SCALAR, 1, 1, false>;
      case 2: return k_bin_sums<Real, LAYOUT_SCALAR, 2, 1, false>;
      case 3: return k_bin_sums<Real, LAYOUT_SCALAR, 3, 1, false>;
      case 4: return k_bin_sums<Real, LAYOUT_SCALAR, 4, 1, false>;
      default: break;
    }
  }
  if (wide_rows(d))
    return wide_by_atom(d, sizeof(Real) == 4 ? BRILLE_FP32 : BRILLE_FP64) ? k_bin_sums<Real, LAYOUT_WIDE, kWideAtom, 0, false>
                                                                         : k_bin_sums<Real, LAYOUT_WIDE, kWideAll, 0, false>;
  return k_bin_sums<Real, LAYOUT_GENERIC, 0, 0, false>;
}

// Mode groups of the sums of a pass of m points (blockIdx.y): wide rows put
// a tile's modes in sequence on its four waves, so a pass of few tiles would
// leave most CUs idle; its modes are split over up to B / 4 blocks per tile
// until the pass has about eight blocks per CU (each group repeats phase 1,
// a few FP64 predicates per point).  Other kernels pick_sums selects (the
// proposal's among them): one group.
static uint32_t sum_groups(const brille_grid* G, uint32_t m, bool prop) {
  if (prop || !wide_rows(G->d)) return 1;
  const uint32_t tiles = (m + TILE - 1) / TILE;
  const uint32_t want = 8u * std::max<uint32_t>(G->n_cu, 1);
  const uint32_t max_groups = std::max<uint32_t>((G->d.B + 3) / 4, 1);
  const uint32_t g = (want + tiles - 1) / std::max<uint32_t>(tiles, 1);
  return std::max<uint32_t>(1, std::min(g, max_groups));
}

// SC of the one-pass sums kernel pick_sums selects, or 0 for the general ones
static int onepass_sc(const DataConst& d) {
  const int L = payload_layout(d);
  if (d.sv != 1) return 0;
  if (L == LAYOUT_GAMMA && (d.sc == 3 || d.sc == 6 || d.sc == 9 || d.sc == 12)) return (int)d.sc;
  if (L == LAYOUT_SCALAR && d.sc >= 1 && d.sc <= 4) return (int)d.sc;
  return 0;
}

template <typename Real>
static uint64_t payload_stride_host(const DataConst& d) {
  return sizeof(Real) == 8 ? d.stride64 : d.stride32;
}

template <typename Real>
static int build_cube_rows(brille_grid* G, hipStream_t s, Real** out) {
  const DataConst& d = G->d;
  const uint64_t cube_units = G->n_cubes * 8 * d.B;
  const uint64_t units = cube_units + G->n_tets * 4 * d.B;
  void* p = nullptr;
  if (hipMalloc(&p, units * payload_stride_host<Real>(d) * sizeof(Real)) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the cube row table");
  G->data_allocs.push_back(p);
  Real* t = static_cast<Real*>(p);
  const dim3 grid((unsigned)((units + 255) / 256));
  switch (onepass_sc(d)) {
    case 1: hipLaunchKernelGGL((k_table_rows<Real, 1>), grid, dim3(256), 0, s, G->g, d, cube_units, units, t); break;
    case 2: hipLaunchKernelGGL((k_table_rows<Real, 2>), grid, dim3(256), 0, s, G->g, d, cube_units, units, t); break;
    case 3: hipLaunchKernelGGL((k_table_rows<Real, 3>), grid, dim3(256), 0, s, G->g, d, cube_units, units, t); break;
    case 4: hipLaunchKernelGGL((k_table_rows<Real, 4>), grid, dim3(256), 0, s, G->g, d, cube_units, units, t); break;
    case 6: hipLaunchKernelGGL((k_table_rows<Real, 6>), grid, dim3(256), 0, s, G->g, d, cube_units, units, t); break;
    case 9: hipLaunchKernelGGL((k_table_rows<Real, 9>), grid, dim3(256), 0, s, G->g, d, cube_units, units, t); break;
    case 12: hipLaunchKernelGGL((k_table_rows<Real, 12>), grid, dim3(256), 0, s, G->g, d, cube_units, units, t); break;
    default: return fail(BRILLE_ERR_UNSUPPORTED, "no cube row kernel for this payload");
  }
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "cube row table launch failed");
  *out = t;
  return BRILLE_OK;
}

template <typename Real>
static int build_wide_factors(brille_grid* G, hipStream_t s, typename Cplx<Real>::T** out) {
  using C = typename Cplx<Real>::T;
  const DataConst& d = G->d;
  const uint64_t cube_units = G->n_cubes * d.B;
  const uint64_t units = cube_units + G->n_tets * d.B;
  void* p = nullptr;
  if (hipMalloc(&p, (G->n_cubes * 8 + G->n_tets * 4) * d.B * sizeof(C)) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the wide factor table");
  G->data_allocs.push_back(p);
  C* t = static_cast<C*>(p);
  hipLaunchKernelGGL((k_wide_factors<Real>), dim3((unsigned)((units + 3) / 4)), dim3(256), 0, s, G->g, d, cube_units,
                     units, t);
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "wide factor table launch failed");
  *out = t;
  return BRILLE_OK;
}

// Builds the payload's derived table of the given precision once (freed with
// the payload): the antiphased rows (8 per cube, 4 per tetrahedron and mode)
// for the one-pass layouts, the antiphase factors alone for wide rows.  The
// general layout, grids without cells and tables beyond kCubeRowsMax go
// without, and their sums compute every antiphase on the fly.
constexpr uint64_t kCubeRowsMax = (4ull << 30) - 1;  // the sums address the table with 32-bit offsets
int ensure_cube_rows(brille_grid* G, int precision, hipStream_t s) {
  DataConst& d = G->d;
  if (G->n_cubes + G->n_tets == 0 || d.B == 0) return BRILLE_OK;
  if (wide_rows(d)) {
    // wide rows: the antiphases only (one complex per slot row)
    const uint64_t fac_bytes = (G->n_cubes * 8 + G->n_tets * 4) * d.B * (precision == BRILLE_FP32 ? 8ull : 16ull);
    if (fac_bytes > kCubeRowsMax) return BRILLE_OK;
    d.tab_tet_row = (uint32_t)(G->n_cubes * 8 * d.B);
    if (precision == BRILLE_FP32) {
      if (d.wide_fac32) return BRILLE_OK;
      float2* t = nullptr;
      const int rc = build_wide_factors<float>(G, s, &t);
      if (rc == BRILLE_OK) d.wide_fac32 = t;
      return rc;
    }
    if (d.wide_fac64) return BRILLE_OK;
    double2* t = nullptr;
    const int rc = build_wide_factors<double>(G, s, &t);
    if (rc == BRILLE_OK) d.wide_fac64 = t;
    return rc;
  }
  if (onepass_sc(d) == 0) return BRILLE_OK;
  const uint64_t row_bytes = precision == BRILLE_FP32 ? d.stride32 * 4ull : d.stride64 * 8ull;
  if ((G->n_cubes * 8 + G->n_tets * 4) * d.B * row_bytes > kCubeRowsMax) return BRILLE_OK;
  d.tab_tet_row = (uint32_t)(G->n_cubes * 8 * d.B);
  if (precision == BRILLE_FP32) {
    if (d.cube_rows32) return BRILLE_OK;
    float* t = nullptr;
    const int rc = build_cube_rows<float>(G, s, &t);
    if (rc == BRILLE_OK) d.cube_rows32 = t;
    return rc;
  }
  if (d.cube_rows64) return BRILLE_OK;
  double* t = nullptr;
  const int rc = build_cube_rows<double>(G, s, &t);
  if (rc == BRILLE_OK) d.cube_rows64 = t;
  return rc;
}

// the aux outputs of a pass that starts at point c0
static brille_aux_out offset_aux(const brille_aux_out& x, uint64_t c0) {
  brille_aux_out o = x;
  if (o.ir_q) o.ir_q += 3 * c0;
  if (o.tau) o.tau += 3 * c0;
  if (o.rot) o.rot += c0;
  if (o.invrot) o.invrot += c0;
  if (o.node) o.node += c0;
  if (o.tet) o.tet += c0;
  if (o.n_iw) o.n_iw += c0;
  if (o.iw_index) o.iw_index += 8 * c0;
  if (o.iw_weight) o.iw_weight += 8 * c0;
  if (o.status) o.status += c0;
  return o;
}

// Buckets of at most kNodesPerBucket trellis nodes: k_bucket_perm orders each
// bucket by node in LDS (kPermNodes counters), so the buckets only have to be
// few enough that a bucketing block (BIN_SPAN points) puts runs of records
// into each: the scatter writes a block's records of one bucket as one run,
// and short runs cost partial-line writes.  Config 4 at 1e8 Q: 1024 nodes per
// bucket 26.7 ms, 2048 26.6 ms, 4096 25.9 ms per call.
constexpr uint32_t kNodesPerBucket = kPermNodes;
static uint32_t bucket_count(const brille_grid* G) {
  const uint64_t n_nodes = std::max<uint64_t>(G->g.n_nodes, 1);
  uint32_t nb = 16;
  while (nb < 2048 && (uint64_t)nb * kNodesPerBucket < n_nodes) nb <<= 1;
  return (uint32_t)std::min<uint64_t>(nb, n_nodes);
}

bool binned_applies(const brille_grid* G, const LaunchArgs& a, int kernel) {
  if (kernel == BRILLE_KERNEL_PLAIN || kernel == BRILLE_KERNEL_TILE) return false;
  const brille_aux_out& x = a.aux;
  // the records carry no tau; the sums keep a Gamma frame per point-group
  // operation (a crystallographic point group has at most 48)
  if (x.tau || G->g.G > kMaxOps) return false;
  if (a.use_proposal || kernel == BRILLE_KERNEL_BINNED) return true;
  if (x.ir_q || x.rot || x.invrot || x.node || x.tet || x.n_iw || x.iw_index || x.iw_weight || x.status)
    return false;
  // automatic: large calls on payloads that do not fit the L2 of one XCD
  const DataConst& d = G->d;
  const uint64_t payload = (uint64_t)d.V * d.B * d.stride32 * 4;
  return a.n >= (1ull << 18) && payload > (8ull << 20);
}

__global__ void k_zero_words(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

void zero_words(void* p, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_zero_words, dim3((n + 255) / 256), dim3(256), 0, s, static_cast<uint32_t*>(p), n);
}

int exclusive_scan_u32(uint32_t* a, uint64_t m, uint32_t* part, hipStream_t s) {
  if (m == 0) return BRILLE_OK;
  const uint64_t np = (m + SCAN_SEG - 1) / SCAN_SEG;
  if (np > 256ull * 1024) return fail(BRILLE_ERR_UNSUPPORTED, "scan too long");
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)np), dim3(BIN_BLOCK), 0, s, a, m, part);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(BIN_BLOCK), 0, s, part, (uint32_t)np);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)np), dim3(BIN_BLOCK), 0, s, a, m, part);
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "scan launch failed");
  return BRILLE_OK;
}

static void free_work_space(brille_grid* G) {
  if (G->bin_ws) (void)hipFree(G->bin_ws);
  G->bin_ws = nullptr;
  G->bin_ws_bytes = 0;
}

void free_binned(brille_grid* G) {
  free_work_space(G);
  for (auto& e : G->bin_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (G->bin_side) (void)hipStreamDestroy(G->bin_side);
  G->bin_side = nullptr;
}

// the side stream and events of the two-stage pipeline, made on first use
// (default priority) and kept with the grid
static int ensure_side_stream(brille_grid* G) {
  if (G->bin_side) return BRILLE_OK;
  for (auto& e : G->bin_ev)
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(BRILLE_ERR_DEVICE, "hipEventCreate failed");
  if (hipStreamCreateWithFlags(&G->bin_side, hipStreamNonBlocking) != hipSuccess) {
    G->bin_side = nullptr;
    return fail(BRILLE_ERR_DEVICE, "hipStreamCreate failed for the binned pipeline");
  }
  return BRILLE_OK;
}

// One half of the work space: the buffers of one pass.
struct PassBuffers {
  FoldRec* rec;
  SortRec* srt;
  uint32_t *hist, *part, *plist, *pcount, *perm, *snode;
};

// A call's passes (BIN_CHUNK points each) run as a two-stage pipeline when
// there are several: the preparation of pass k + 1 (fold, scans, scatter,
// node order, proposal) on the grid's side stream while the sums of pass k run
// on the call's stream, each pass with its own half of the work space.  The
// fold is bound by FP64 issue and the sums by their scattered stores, and one
// fold wave per SIMD fits beside four sums waves, so the two overlap.  The
// call stream waits for each preparation before its sums, the side stream for
// the sums that last used a half, and the side stream starts after what the
// call stream held before the call: the call is ordered on `s` like a single
// kernel.
//
// stage_ms (nullable): adds the milliseconds of fold, bucket (scans, scatter,
// node order, proposal) and sums over all passes, each from events on the
// stream that ran it (with several passes the stages overlap, so they add up
// to more than the call)
// Byte offsets of one pass's buffers (a half of the work space) for passes of
// `chunk` points.
struct PassLayout {
  uint64_t rec, srt, hist, part, list, perm, node, cnt, half;
  uint64_t scan_blocks;  // segments of the histogram scan (k_scan_top handles at most 256 Ki)
};
static PassLayout pass_layout(uint64_t chunk, uint32_t nb) {
  const uint32_t nblk_max = (uint32_t)((chunk + BIN_SPAN - 1) / BIN_SPAN);
  const uint64_t m_max = (uint64_t)(nb + 1) * nblk_max;
  const uint64_t np_max = (m_max + SCAN_SEG - 1) / SCAN_SEG;
  auto al = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
  PassLayout L;
  L.rec = 0;
  L.srt = al(chunk * sizeof(FoldRec));
  L.hist = L.srt + al(chunk * sizeof(SortRec));
  L.part = L.hist + al(m_max * 4);
  L.list = L.part + al(np_max * 4);
  L.perm = L.list + al(chunk * 4);
  L.node = L.perm + al(chunk * 4);
  L.cnt = L.node + al(chunk * 4);
  L.half = L.cnt + 256;
  L.scan_blocks = np_max;
  return L;
}

uint64_t binned_pass_bytes(const brille_grid* G, uint64_t chunk) { return pass_layout(chunk, bucket_count(G)).half; }

// The largest pass (a power of two, at most BIN_CHUNK points) whose buffers fit `bytes`.
uint64_t binned_pass_for(const brille_grid* G, uint64_t bytes) {
  uint64_t c = BIN_CHUNK;
  while (c > BIN_SPAN && binned_pass_bytes(G, c) > bytes) c >>= 1;
  return binned_pass_bytes(G, c) <= bytes ? c : 0;
}

int launch_binned(brille_grid* G, const LaunchArgs& a, int precision, hipStream_t s, float* stage_ms) {
  if (int rc = ensure_cube_rows(G, precision, s); rc != BRILLE_OK) return rc;
  const uint64_t chunk = std::min<uint64_t>(a.n, BIN_CHUNK);
  if (chunk == 0) return BRILLE_OK;
  const uint64_t passes = (a.n + chunk - 1) / chunk;
  const int halves = passes > 1 ? 2 : 1;
  const uint64_t need = binned_pass_bytes(G, chunk) * (uint64_t)halves;
  if (G->bin_ws_bytes < need) {
    free_work_space(G);
    if (hipMalloc(&G->bin_ws, need) != hipSuccess) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the binned work space");
    G->bin_ws_bytes = need;
  }
  if (halves > 1)
    if (int rc = ensure_side_stream(G); rc != BRILLE_OK) return rc;
  return run_binned_passes(G, a, precision, s, stage_ms, G->bin_ws, chunk, halves);
}

// The pipeline over work space `ws` in passes of `chunk` points; with two
// halves the preparation of the next pass overlaps the sums (side stream
// and events of the grid), with one everything runs on `s` in order.
int run_binned_passes(brille_grid* G, const LaunchArgs& a, int precision, hipStream_t s, float* stage_ms, void* ws_base,
                      uint64_t chunk, int halves) {
  if (a.n == 0) return BRILLE_OK;
  const uint64_t passes = (a.n + chunk - 1) / chunk;
  const uint32_t nb = bucket_count(G);
  const PassLayout Lo = pass_layout(chunk, nb);
  if (Lo.scan_blocks > 256ull * 1024) return fail(BRILLE_ERR_INVALID_ARGUMENT, "binned pass too large");
  if (G->g.n_nodes >= kNodePoly) return fail(BRILLE_ERR_INVALID_ARGUMENT, "the binned kernel takes grids of fewer than 2^31 nodes");
  const uint64_t o_rec = Lo.rec, o_srt = Lo.srt, o_hist = Lo.hist, o_part = Lo.part, o_list = Lo.list,
                 o_perm = Lo.perm, o_node = Lo.node, o_cnt = Lo.cnt, half = Lo.half;
  const bool prop = a.use_proposal != 0;
  PassBuffers buf[2];
  for (int h = 0; h < halves; ++h) {
    char* ws = static_cast<char*>(ws_base) + h * half;
    PassBuffers& b = buf[h];
    b.rec = reinterpret_cast<FoldRec*>(ws + o_rec);
    b.srt = reinterpret_cast<SortRec*>(ws + o_srt);
    b.hist = reinterpret_cast<uint32_t*>(ws + o_hist);
    b.part = reinterpret_cast<uint32_t*>(ws + o_part);
    b.plist = reinterpret_cast<uint32_t*>(ws + o_list);
    b.pcount = reinterpret_cast<uint32_t*>(ws + o_cnt);
    b.perm = reinterpret_cast<uint32_t*>(ws + o_perm);
    b.snode = reinterpret_cast<uint32_t*>(ws + o_node);
  }
  // timing: five events per pass (fold start, fold end, bucket end on the
  // preparing stream; sums start and end on the call stream)
  std::vector<hipEvent_t> tev;
  auto destroy_tev = [&tev]() {
    for (auto e : tev)
      if (e) (void)hipEventDestroy(e);
  };
  if (stage_ms) {
    tev.assign(passes * 5, nullptr);
    for (auto& e : tev)
      if (hipEventCreate(&e) != hipSuccess) {
        destroy_tev();
        return fail(BRILLE_ERR_DEVICE, "hipEventCreate failed");
      }
  }
  auto mark = [&](uint64_t k, int j, hipStream_t st) {
    if (stage_ms) (void)hipEventRecord(tev[k * 5 + j], st);
  };
  const SumsKernel sums = precision == BRILLE_FP32 ? pick_sums<float>(G->d, prop) : pick_sums<double>(G->d, prop);
  const bool by_atom = !prop && wide_rows(G->d) && wide_by_atom(G->d, precision);
  const SumsKernel rest_sums = precision == BRILLE_FP32 ? k_bin_sums<float, LAYOUT_WIDE, kWideRest, 0, false>
                                                        : k_bin_sums<double, LAYOUT_WIDE, kWideRest, 0, false>;
  const size_t rs = precision == BRILLE_FP32 ? 4 : 8;
  const size_t vb = (size_t)G->d.B * G->d.sv * rs, cb = (size_t)G->d.B * G->d.sc * 2 * rs;
  const size_t lds1 = (size_t)(nb + 1) * 4;
  hipStream_t side = halves > 1 ? G->bin_side : s;
  enum { EV_IN = 0, EV_PREP = 1, EV_DONE = 3 };  // G->bin_ev: call start, prepared[2], summed[2]
  if (halves > 1) {
    (void)hipEventRecord(G->bin_ev[EV_IN], s);
    (void)hipStreamWaitEvent(side, G->bin_ev[EV_IN], 0);
  }
  for (uint64_t k = 0; k < passes; ++k) {
    const uint64_t c0 = k * chunk;
    const int h = (int)(k & 1u) % halves;
    const PassBuffers& b = buf[h];
    const uint32_t m = (uint32_t)std::min<uint64_t>(chunk, a.n - c0);
    const uint32_t nblk = (m + BIN_SPAN - 1) / BIN_SPAN;
    const uint64_t M = (uint64_t)(nb + 1) * nblk;
    const uint32_t np = (uint32_t)((M + SCAN_SEG - 1) / SCAN_SEG);
    // preparation of pass k on the side stream, once the sums of pass k - 2
    // no longer read this half
    if (halves > 1 && k >= 2) (void)hipStreamWaitEvent(side, G->bin_ev[EV_DONE + h], 0);
    if (prop) zero_words(b.pcount, 1, side);
    mark(k, 0, side);
    hipLaunchKernelGGL(k_bin_fold<true>, dim3(nblk), dim3(BIN_BLOCK), lds1, side, G->g, a.Q + 3 * c0, m, a.fold,
                       b.rec, b.hist, nblk, nb, a.err, a.base + c0, prop ? 1 : 0);
    mark(k, 1, side);
    hipLaunchKernelGGL(k_scan_reduce, dim3(np), dim3(BIN_BLOCK), 0, side, b.hist, M, b.part);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(BIN_BLOCK), 0, side, b.part, np);
    hipLaunchKernelGGL(k_scan_apply, dim3(np), dim3(BIN_BLOCK), 0, side, b.hist, M, b.part);
    hipLaunchKernelGGL(k_bin_scatter, dim3(nblk), dim3(BIN_BLOCK), 2 * lds1, side, b.rec, m, b.hist, nblk, nb,
                       G->g.n_nodes, b.srt, b.snode);
    hipLaunchKernelGGL(k_bucket_perm, dim3(nb + 1), dim3(PERM_BLOCK), 0, side, b.snode, m, b.hist, nblk, nb,
                       G->g.n_nodes, b.perm);
    if (prop) {
      hipLaunchKernelGGL(k_poly_list, dim3((m + LIST_SPAN - 1) / LIST_SPAN), dim3(BIN_BLOCK), 0, side, b.perm, m,
                         b.plist, b.pcount);
      hipLaunchKernelGGL(k_bin_propose, dim3(std::min<uint32_t>((m + 63) / 64, std::max<uint32_t>(G->n_cu, 1) * 4)),
                         dim3(256), 0, side, G->g, a.prop, b.srt, b.plist, b.pcount);
    }
    mark(k, 2, side);
    if (halves > 1) {
      (void)hipEventRecord(G->bin_ev[EV_PREP + h], side);
      (void)hipStreamWaitEvent(s, G->bin_ev[EV_PREP + h], 0);
    }
    // sums of pass k on the call stream
    mark(k, 3, s);
    const dim3 sgrid((m + TILE - 1) / TILE, sum_groups(G, m, prop));
    if (by_atom) zero_words(b.pcount, 1, s);
    hipLaunchKernelGGL(sums, sgrid, dim3(TILE), 0, s, G->g, G->d, a.prop, b.srt, b.perm, m,
                       static_cast<char*>(a.vals) + c0 * vb, static_cast<char*>(a.vecs) + c0 * cb,
                       offset_aux(a.aux, c0), a.err, a.base + c0, b.plist, b.pcount);
    if (by_atom)  // the listed points without a table block: one block per CU at most
      hipLaunchKernelGGL(rest_sums, dim3(std::min<uint32_t>(sgrid.x, std::max<uint32_t>(G->n_cu, 1)), sgrid.y),
                         dim3(TILE), 0, s, G->g, G->d, a.prop, b.srt, b.perm, m,
                         static_cast<char*>(a.vals) + c0 * vb, static_cast<char*>(a.vecs) + c0 * cb,
                         offset_aux(a.aux, c0), a.err, a.base + c0, b.plist, b.pcount);
    mark(k, 4, s);
    if (halves > 1) (void)hipEventRecord(G->bin_ev[EV_DONE + h], s);
  }
  if (stage_ms) {
    (void)hipEventSynchronize(tev[(passes - 1) * 5 + 4]);
    for (uint64_t k = 0; k < passes; ++k) {
      const int from[3] = {0, 1, 3}, to[3] = {1, 2, 4};
      for (int j = 0; j < 3; ++j) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, tev[k * 5 + from[j]], tev[k * 5 + to[j]]);
        stage_ms[j] += ms;
      }
    }
    destroy_tev();
  }
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "binned launch failed");
  return BRILLE_OK;
}

}  // namespace brille_amd
