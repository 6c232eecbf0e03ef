// Learned tetrahedron proposal: training on the GPU and the C API (no
// reference counterpart; the network only proposes, the FP64 predicates of
// the reference decide -- see proposal.hpp).
//
// Training pairs come from the exact locator: random Q, uniform in
// [-2, 2]^3 r.l.u., folded into the irreducible wedge (stage (a)), located in
// the trellis; the points that land in polyhedron nodes give (features,
// tetrahedron of PolyNode::indices_weights).  The network is trained with
// softmax cross-entropy restricted to the tetrahedra of the point's node, by
// this file's own kernels; the three 128-wide products of a step run on the
// matrix cores in bf16 with FP32 accumulation (v_mfma_f32_16x16x32_bf16),
// everything else in FP32:
//   k_tr_layer1      h1 = relu(W1 x + b1) per sample (FP32, and a bf16 copy);
//   k_tr_layer2      a2 = bf16(h1) bf16(W2)^T + b2, one wave per 16 samples;
//   k_tr_head        one wave per sample: h2 = relu(a2), the logits of the
//                    node's candidates (FP32 W3), loss, the sparse W3/b3
//                    gradients (atomics on the candidate rows) and
//                    da2 = (dz W3) * [a2 > 0] (FP32, and a bf16 copy);
//   k_tr_back1       da1 = (bf16(da2) bf16(W2)) * [h1 > 0];
//   k_tr_wgrad2      gW2 = bf16(da2)^T bf16(h1) over the batch (K = samples);
//   k_tr_bias2, k_prop_wgrad1   gb2, gW1 and gb1 as FP32 sums over the batch;
//   k_prop_adam      Adam on the FP32 master parameters (and clears gradients).
// After training the inference weights are exported as bf16 (W3 lane-major)
// and the hit rate is measured on held-out pairs with the matrix-core
// inference the tile kernels use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "grid_state.hpp"

using namespace brille_amd;

namespace brille_amd {

constexpr int H = kPropH;
constexpr uint32_t kMaxCand = 64;     // tetrahedra per polyhedron node the trainer handles
constexpr double kPropMargin = 1e-6;  // barycentric depth a proposal must show to be taken

struct ParamOffsets {
  uint64_t W1, b1, W2, b2, W3, b3, n;
};
__host__ __device__ inline ParamOffsets param_offsets(uint32_t K) {
  ParamOffsets o;
  o.W1 = 0;
  o.b1 = o.W1 + 3 * H;
  o.W2 = o.b1 + H;
  o.b2 = o.W2 + (uint64_t)H * H;
  o.W3 = o.b2 + H;
  o.b3 = o.W3 + (uint64_t)K * H;
  o.n = o.b3 + K;
  return o;
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double unit_double(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// labelled pairs from random Q: features, tetrahedron, candidates of its node
__global__ __launch_bounds__(256) void k_prop_label(GridConst g, ProposalConst pc, uint64_t seed, uint64_t start,
                                                    uint32_t m, uint32_t cap, uint32_t* count, float* feat,
                                                    int32_t* label, uint32_t* t0, uint32_t* nc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t id = start + i;
  double Q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) Q[d] = 4.0 * unit_double(splitmix64(seed * 0x632BE59BD9B4E019ull + 3 * id + d)) - 2.0;
  Located L;
  fold_node<true>(g, Q, FOLD_IR, L);
  if (L.status != BRILLE_PT_OK || g.node_type[L.node] != BRILLE_NODE_POLY) return;
  node_slots(g, L);
  if (L.status != BRILLE_PT_OK || L.tet == kNoTet) return;
  const uint32_t p = g.node_index[L.node];
  const uint32_t a = g.poly_off[p], b = g.poly_off[p + 1];
  if (b - a > kMaxCand) return;
  const uint32_t slot = atomicAdd(count, 1u);
  if (slot >= cap) return;
  double x[3];
  mat_vec(x, g.outer_to_xyz, L.q);
  float f[3];
  prop_features(g, x, L.node, f);
  feat[3 * (uint64_t)slot] = f[0];
  feat[3 * (uint64_t)slot + 1] = f[1];
  feat[3 * (uint64_t)slot + 2] = f[2];
  label[slot] = (int32_t)L.tet;
  t0[slot] = a;
  nc[slot] = b - a;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// gW1[k][d] += sum_s DA1[s][k] x[s][d], gb1[k] += sum_s DA1[s][k]: block (segment), thread k
__global__ __launch_bounds__(128) void k_prop_wgrad1(const float* DA1, const float* feat, uint32_t n, uint32_t seg,
                                                     float* G, uint64_t oW1, uint64_t ob1) {
  const uint32_t k = threadIdx.x;
  const uint32_t s0 = blockIdx.x * seg, s1 = min(n, s0 + seg);
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, ab = 0.f;
  for (uint32_t s = s0; s < s1; ++s) {
    const float d = DA1[(uint64_t)s * H + k];
    a0 += d * feat[3 * (uint64_t)s];
    a1 += d * feat[3 * (uint64_t)s + 1];
    a2 += d * feat[3 * (uint64_t)s + 2];
    ab += d;
  }
  atomicAdd(&G[oW1 + 3 * k], a0);
  atomicAdd(&G[oW1 + 3 * k + 1], a1);
  atomicAdd(&G[oW1 + 3 * k + 2], a2);
  atomicAdd(&G[ob1 + k], ab);
}

__global__ __launch_bounds__(256) void k_prop_adam(float* P, float* G, float* M, float* V, uint64_t n, float lr,
                                                   float c1, float c2) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = 0.9f * M[i] + 0.1f * g;
  const float v = 0.999f * V[i] + 0.001f * g * g;
  M[i] = m;
  V[i] = v;
  P[i] -= lr * (m * c1) / (sqrtf(v * c2) + 1e-8f);
  G[i] = 0.f;
}

__global__ void k_fill_zero(float* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite inputs)
}

// ---- training kernels (bf16 matrix cores for the 128-wide products) ----

// bf16 copies of W2 for one step: row-major (rows = output unit) and transposed
__global__ void k_tr_w2(const float* P, uint64_t oW2, uint16_t* W2b, uint16_t* W2bT) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)H * H) return;
  const uint32_t j = i / H, k = i % H;
  const uint16_t v = bf16_bits(P[oW2 + i]);
  W2b[i] = v;
  W2bT[k * H + j] = v;
}

// h1 = relu(W1 x + b1) for samples [0, n_pad): FP32 and bf16 rows; samples
// from n on are zero (the products then see zero rows)
__global__ __launch_bounds__(256) void k_tr_layer1(const float* P, uint64_t oW1, uint64_t ob1, const float* feat,
                                                   uint32_t n, uint32_t n_pad, float* H1, uint16_t* H1b) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n_pad * H) return;
  const uint32_t s = (uint32_t)(t / H), k = (uint32_t)(t % H);
  float h = 0.f;
  if (s < n) {
    const float a = P[oW1 + 3 * k] * feat[3 * (uint64_t)s] + P[oW1 + 3 * k + 1] * feat[3 * (uint64_t)s + 1] +
                    P[oW1 + 3 * k + 2] * feat[3 * (uint64_t)s + 2] + P[ob1 + k];
    h = fmaxf(a, 0.f);
  }
  H1[t] = h;
  H1b[t] = bf16_bits(h);
}

// out[s][c] = sum_k A[s][k] B[c][k] (+ bias[c]) for 16 samples per wave: A
// rows bf16 (samples x 128), B bf16 row-major (128 x 128, row = output c);
// out FP32 (samples x 128).  With mask (nullable, FP32 samples x 128) the
// product is kept only where mask > 0.
__global__ __launch_bounds__(256) void k_tr_rows_x_w(const uint16_t* A, const uint16_t* Bw, const float* bias,
                                                     const float* mask, uint32_t n_pad, float* out) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t s0 = wave * 16;
  if (s0 >= n_pad) return;  // whole waves leave together
  const uint32_t l = threadIdx.x & 63u, m = l & 15u, grp = l >> 4;
  bf16x8_t a[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) a[kb] = load_bf16x8(A + (uint64_t)(s0 + m) * H + 32 * kb + 8 * grp);
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    floatx4_t acc = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kb], load_bf16x8(Bw + (uint64_t)(nb * 16 + m) * H + 32 * kb + 8 * grp),
                                                    acc, 0, 0, 0);
    const uint32_t c = nb * 16 + m;
    const float bc = bias ? bias[c] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint64_t at = (uint64_t)(s0 + 4 * grp + r) * H + c;
      float v = acc[r] + bc;
      if (mask && !(mask[at] > 0.f)) v = 0.f;
      out[at] = v;
    }
  }
}

// one wave per sample: the head (FP32), its loss and sparse gradients, and
// da2 = (dz W3) * [a2 > 0] as FP32 and bf16 rows (zero rows from n to n_pad)
__global__ __launch_bounds__(256) void k_tr_head(const float* P, float* G, uint32_t K, const float* A2,
                                                 const int32_t* label, const uint32_t* t0, const uint32_t* nc,
                                                 uint32_t n, uint32_t n_pad, float inv_n, float* DA2, uint16_t* DA2b,
                                                 float* loss) {
  const ParamOffsets o = param_offsets(K);
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
  const uint32_t s = blockIdx.x * 4 + w;
  if (s >= n_pad) return;
  const uint32_t j0 = l, j1 = l + 64;
  const float a2_0 = A2[(uint64_t)s * H + j0], a2_1 = A2[(uint64_t)s * H + j1];
  const float h2_0 = fmaxf(a2_0, 0.f), h2_1 = fmaxf(a2_1, 0.f);
  float dh2_0 = 0.f, dh2_1 = 0.f;
  if (s < n) {
    const uint32_t c0 = t0[s], cn = nc[s];
    const int32_t lab = label[s] - (int32_t)c0;
    const float* W3 = P + o.W3;
    float z = -3.0e38f;
    for (uint32_t c = 0; c < cn; ++c) {
      const uint64_t row = (uint64_t)(c0 + c) * H;
      float part = W3[row + j0] * h2_0 + W3[row + j1] * h2_1;
      part = wave_sum(part) + P[o.b3 + c0 + c];
      if (l == c) z = part;
    }
    const float mx = wave_max(l < cn ? z : -3.0e38f);
    const float e = l < cn ? expf(z - mx) : 0.f;
    const float S = wave_sum(e);
    const float zl = __shfl(z, lab);
    if (l == 0) atomicAdd(loss, -(zl - mx - logf(S)) * inv_n);
    float dz = l < cn ? e / S : 0.f;
    if ((int32_t)l == lab) dz -= 1.f;
    dz *= inv_n;
    for (uint32_t c = 0; c < cn; ++c) {
      const float dzc = __shfl(dz, (int)c);
      const uint64_t row = (uint64_t)(c0 + c) * H;
      atomicAdd(&G[o.W3 + row + j0], dzc * h2_0);
      atomicAdd(&G[o.W3 + row + j1], dzc * h2_1);
      dh2_0 += dzc * W3[row + j0];
      dh2_1 += dzc * W3[row + j1];
    }
    if (l < cn) atomicAdd(&G[o.b3 + c0 + l], dz);
  }
  const float da2_0 = a2_0 > 0.f ? dh2_0 : 0.f, da2_1 = a2_1 > 0.f ? dh2_1 : 0.f;
  DA2[(uint64_t)s * H + j0] = da2_0;
  DA2[(uint64_t)s * H + j1] = da2_1;
  DA2b[(uint64_t)s * H + j0] = bf16_bits(da2_0);
  DA2b[(uint64_t)s * H + j1] = bf16_bits(da2_1);
}

// bf16 transpose of a (rows x 128) matrix into (128 x rows), 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void k_tr_transpose(const uint16_t* in, uint32_t rows, uint16_t* out) {
  __shared__ uint16_t t[32][33];
  const uint32_t r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;  // 32 x 8
  for (uint32_t y = ty; y < 32; y += 8) t[y][tx] = in[(uint64_t)(r0 + y) * H + c0 + tx];
  __syncthreads();
  for (uint32_t y = ty; y < 32; y += 8) out[(uint64_t)(c0 + y) * rows + r0 + tx] = t[tx][y];
}

// gW2[j][k] += sum over the chunk's samples of DA2T[j][s] H1T[k][s]: one
// wave per (16 rows j, chunk of samples), all 128 columns; FP32 atomics
constexpr uint32_t kTrChunk = 1024;
__global__ __launch_bounds__(64) void k_tr_wgrad2(const uint16_t* DA2T, const uint16_t* H1T, uint32_t n_pad, float* G,
                                                  uint64_t oW2) {
  const uint32_t jt = blockIdx.x, chunk = blockIdx.y;
  const uint32_t l = threadIdx.x, m = l & 15u, grp = l >> 4;
  const uint32_t s_lo = chunk * kTrChunk, s_hi = min(n_pad, s_lo + kTrChunk);
  floatx4_t acc[8];
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) acc[nb] = floatx4_t{0.f, 0.f, 0.f, 0.f};
  for (uint32_t s = s_lo; s < s_hi; s += 32) {
    const bf16x8_t a = load_bf16x8(DA2T + (uint64_t)(jt * 16 + m) * n_pad + s + 8 * grp);
#pragma unroll
    for (int nb = 0; nb < 8; ++nb)
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, load_bf16x8(H1T + (uint64_t)(nb * 16 + m) * n_pad + s + 8 * grp),
                                                        acc[nb], 0, 0, 0);
  }
#pragma unroll
  for (int nb = 0; nb < 8; ++nb)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      atomicAdd(&G[oW2 + (uint64_t)(jt * 16 + 4 * grp + r) * H + nb * 16 + m], acc[nb][r]);
}

// gb2[j] += sum_s DA2[s][j]: block = segment of samples, thread j
__global__ __launch_bounds__(128) void k_tr_bias2(const float* DA2, uint32_t n, uint32_t seg, float* G, uint64_t ob2) {
  const uint32_t j = threadIdx.x;
  const uint32_t s0 = blockIdx.x * seg, s1 = min(n, s0 + seg);
  float acc = 0.f;
  for (uint32_t s = s0; s < s1; ++s) acc += DA2[(uint64_t)s * H + j];
  atomicAdd(&G[ob2 + j], acc);
}

// bf16 inference copies: W2 row-major, W3 lane-major (element nb*16+m of a row at m*8+nb)
__global__ void k_prop_export(const float* P, uint32_t K, uint16_t* W2b, uint16_t* W3b) {
  const ParamOffsets o = param_offsets(K);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (uint64_t)H * H) W2b[i] = bf16_bits(P[o.W2 + i]);
  if (i < (uint64_t)K * H) {
    const uint64_t c = i / H;
    const uint32_t e = (uint32_t)(i % H), nb = e / 16, m = e % 16;
    W3b[c * H + m * 8 + nb] = bf16_bits(P[o.W3 + i]);
  }
}

// proposals for a list of points (features, candidates): the production inference
__global__ __launch_bounds__(256) void k_prop_predict(ProposalConst pc, const float* feat, const uint32_t* t0,
                                                      const uint32_t* nc, uint32_t n, uint32_t* out) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t c0 = (uint64_t)wave * 16;
  if (c0 >= n) return;  // whole waves leave together
  const uint32_t k = n - c0 < 16 ? (uint32_t)(n - c0) : 16u;
  const PropWeights<const float*, const uint16_t*> W{pc.W1, pc.b1, pc.b2, pc.W2};
  propose16(pc, W, feat + 3 * c0, t0 + c0, nc + c0, k, out + c0);
}

__global__ void k_count_equal(const uint32_t* a, const int32_t* b, uint32_t n, unsigned int* hits) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && a[i] == (uint32_t)b[i]) atomicAdd(hits, 1u);
}

#define P_TRY(expr)                                                                                 \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return fail(BRILLE_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int dalloc(brille_grid* G, T** p, uint64_t n) {
  void* q = nullptr;
  P_TRY(hipMalloc(&q, std::max<uint64_t>(n, 1) * sizeof(T)));
  G->prop_allocs.push_back(q);
  *p = static_cast<T*>(q);
  return BRILLE_OK;
}

void free_proposal(brille_grid* G) {
  for (void* p : G->prop_allocs) (void)hipFree(p);
  G->prop_allocs.clear();
  G->prop_trained = false;
  G->prop_params = G->prop_m = G->prop_v = G->prop_grads = nullptr;
  G->prop_stats = nullptr;
  G->prop_n_params = 0;
  G->prop_step = 0;
  std::memset(&G->prop, 0, sizeof(G->prop));
}

// parameter, moment and inference buffers, He-initialised parameters
static int prop_setup(brille_grid* G, uint64_t seed) {
  if (G->n_polys == 0) return fail(BRILLE_ERR_INVALID_ARGUMENT, "the trellis has no polyhedron nodes to propose for");
  free_proposal(G);
  const uint32_t K = (uint32_t)G->n_tets;
  const ParamOffsets o = param_offsets(K);
  G->prop_K = K;
  G->prop_n_params = o.n;
  int rc = dalloc(G, &G->prop_params, o.n);
  if (rc == BRILLE_OK) rc = dalloc(G, &G->prop_m, o.n);
  if (rc == BRILLE_OK) rc = dalloc(G, &G->prop_v, o.n);
  if (rc == BRILLE_OK) rc = dalloc(G, &G->prop_grads, o.n);
  uint16_t *W2b = nullptr, *W3b = nullptr;
  if (rc == BRILLE_OK) rc = dalloc(G, &W2b, (uint64_t)H * H);
  if (rc == BRILLE_OK) rc = dalloc(G, &W3b, (uint64_t)K * H);
  if (rc == BRILLE_OK) rc = dalloc(G, &G->prop_stats, 2);
  if (rc != BRILLE_OK) return rc;
  std::vector<float> host(o.n, 0.f);
  std::mt19937_64 gen(seed ^ 0xA24BAED4963EE407ull);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (uint64_t i = 0; i < 3 * (uint64_t)H; ++i) host[o.W1 + i] = nd(gen) * std::sqrt(2.f / 3.f) * 2.f;
  for (uint64_t i = 0; i < (uint64_t)H; ++i) host[o.b1 + i] = 0.5f * (nd(gen));
  for (uint64_t i = 0; i < (uint64_t)H * H; ++i) host[o.W2 + i] = nd(gen) * std::sqrt(2.f / H);
  for (uint64_t i = 0; i < (uint64_t)K * H; ++i) host[o.W3 + i] = nd(gen) * std::sqrt(1.f / H);
  P_TRY(hipMemcpy(G->prop_params, host.data(), o.n * sizeof(float), hipMemcpyHostToDevice));
  const unsigned nb = (unsigned)((o.n + 255) / 256);
  hipLaunchKernelGGL(k_fill_zero, dim3(nb), dim3(256), 0, 0, G->prop_m, o.n);
  hipLaunchKernelGGL(k_fill_zero, dim3(nb), dim3(256), 0, 0, G->prop_v, o.n);
  hipLaunchKernelGGL(k_fill_zero, dim3(nb), dim3(256), 0, 0, G->prop_grads, o.n);
  P_TRY(hipGetLastError());
  ProposalConst& pc = G->prop;
  pc.K = K;
  pc.W1 = G->prop_params + o.W1;
  pc.b1 = G->prop_params + o.b1;
  pc.W2 = W2b;
  pc.b2 = G->prop_params + o.b2;
  pc.W3 = W3b;
  pc.b3 = G->prop_params + o.b3;
  pc.margin = kPropMargin;
  pc.stats = G->prop_stats;
  pc.enabled = 0;
  G->prop_step = 0;
  return BRILLE_OK;
}

static int prop_export(brille_grid* G) {
  const uint64_t m = std::max<uint64_t>((uint64_t)H * H, (uint64_t)G->prop_K * H);
  hipLaunchKernelGGL(k_prop_export, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, 0, G->prop_params, G->prop_K,
                     const_cast<uint16_t*>(G->prop.W2), const_cast<uint16_t*>(G->prop.W3));
  P_TRY(hipGetLastError());
  P_TRY(hipDeviceSynchronize());
  G->prop.enabled = 1;
  G->prop_trained = true;
  return BRILLE_OK;
}

// candidates of each label: the polyhedron node whose tetrahedra hold it
static int host_candidates(const brille_grid* G, const int32_t* labels, uint32_t n, std::vector<uint32_t>& t0,
                           std::vector<uint32_t>& nc) {
  const auto& off = G->host_desc_poly_off;
  t0.resize(n);
  nc.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (labels[i] < 0 || (uint64_t)labels[i] >= G->n_tets) return fail(BRILLE_ERR_INVALID_ARGUMENT, "label out of range");
    const auto it = std::upper_bound(off.begin(), off.end(), (uint32_t)labels[i]);
    const size_t p = (size_t)(it - off.begin()) - 1;
    t0[i] = off[p];
    nc[i] = off[p + 1] - off[p];
    if (nc[i] > kMaxCand) return fail(BRILLE_ERR_UNSUPPORTED, "polyhedron node with more than 64 tetrahedra");
  }
  return BRILLE_OK;
}

// one optimisation step's gradients (accumulated into prop_grads) and loss;
// buffers for batches of up to `cap` samples (rounded up to 32)
struct StepBuffers {
  float *H1, *A2, *DA2, *DA1, *loss;
  uint16_t *H1b, *H1T, *DA2b, *DA2T, *W2b, *W2bT;
};
static uint32_t step_pad(uint32_t n) { return (n + 31u) & ~31u; }
static bool step_alloc(StepBuffers& sb, uint32_t cap, const std::function<void*(size_t)>& alloc) {
  const size_t rows = step_pad(cap);
  sb.H1 = static_cast<float*>(alloc(rows * H * 4));
  sb.A2 = static_cast<float*>(alloc(rows * H * 4));
  sb.DA2 = static_cast<float*>(alloc(rows * H * 4));
  sb.DA1 = static_cast<float*>(alloc(rows * H * 4));
  sb.loss = static_cast<float*>(alloc(4));
  sb.H1b = static_cast<uint16_t*>(alloc(rows * H * 2));
  sb.H1T = static_cast<uint16_t*>(alloc(rows * H * 2));
  sb.DA2b = static_cast<uint16_t*>(alloc(rows * H * 2));
  sb.DA2T = static_cast<uint16_t*>(alloc(rows * H * 2));
  sb.W2b = static_cast<uint16_t*>(alloc(H * H * 2));
  sb.W2bT = static_cast<uint16_t*>(alloc(H * H * 2));
  return sb.H1 && sb.A2 && sb.DA2 && sb.DA1 && sb.loss && sb.H1b && sb.H1T && sb.DA2b && sb.DA2T && sb.W2b && sb.W2bT;
}
static void grads_launch(brille_grid* G, const StepBuffers& sb, const float* feat, const int32_t* label,
                         const uint32_t* t0, const uint32_t* nc, uint32_t n, hipStream_t s) {
  const ParamOffsets o = param_offsets(G->prop_K);
  const uint32_t np = step_pad(n);
  const float* P = G->prop_params;
  hipLaunchKernelGGL(k_tr_w2, dim3(H * H / 256), dim3(256), 0, s, P, o.W2, sb.W2b, sb.W2bT);
  hipLaunchKernelGGL(k_tr_layer1, dim3((unsigned)(((uint64_t)np * H + 255) / 256)), dim3(256), 0, s, P, o.W1, o.b1, feat,
                     n, np, sb.H1, sb.H1b);
  hipLaunchKernelGGL(k_tr_rows_x_w, dim3((np / 16 + 3) / 4), dim3(256), 0, s, sb.H1b, sb.W2b, P + o.b2,
                     (const float*)nullptr, np, sb.A2);
  hipLaunchKernelGGL(k_tr_head, dim3((np + 3) / 4), dim3(256), 0, s, P, G->prop_grads, G->prop_K, sb.A2, label, t0, nc,
                     n, np, 1.f / (float)n, sb.DA2, sb.DA2b, sb.loss);
  hipLaunchKernelGGL(k_tr_rows_x_w, dim3((np / 16 + 3) / 4), dim3(256), 0, s, sb.DA2b, sb.W2bT, (const float*)nullptr,
                     sb.H1, np, sb.DA1);
  hipLaunchKernelGGL(k_tr_transpose, dim3(np / 32, H / 32), dim3(256), 0, s, sb.H1b, np, sb.H1T);
  hipLaunchKernelGGL(k_tr_transpose, dim3(np / 32, H / 32), dim3(256), 0, s, sb.DA2b, np, sb.DA2T);
  hipLaunchKernelGGL(k_tr_wgrad2, dim3(H / 16, (np + kTrChunk - 1) / kTrChunk), dim3(64), 0, s, sb.DA2T, sb.H1T, np,
                     G->prop_grads, o.W2);
  // short segments: each thread's sum is a handful of independent loads, and
  // the blocks spread over the chip (512-sample segments left 16 blocks)
  const uint32_t seg = 32, nseg = (n + seg - 1) / seg;
  hipLaunchKernelGGL(k_tr_bias2, dim3(nseg), dim3(H), 0, s, sb.DA2, n, seg, G->prop_grads, o.b2);
  hipLaunchKernelGGL(k_prop_wgrad1, dim3(nseg), dim3(H), 0, s, sb.DA1, feat, n, seg, G->prop_grads, o.W1, o.b1);
}

}  // namespace brille_amd

extern "C" {

int brille_proposal_train(brille_grid* G, const brille_train_opts* opts_in, brille_train_report* report) {
  BRILLE_LOCK_GRID(G);
  if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
  brille_train_opts o{};
  if (opts_in) o = *opts_in;
  if (o.n_train == 0) o.n_train = 20000000ull;
  if (o.epochs == 0) o.epochs = 8;
  if (o.batch == 0) o.batch = 8192;
  if (!(o.learning_rate > 0.f)) o.learning_rate = 2e-3f;
  if (o.n_train > (1ull << 31)) return fail(BRILLE_ERR_INVALID_ARGUMENT, "n_train too large");
  const auto t_start = std::chrono::steady_clock::now();
  int rc = prop_setup(G, o.seed);
  if (rc != BRILLE_OK) return rc;
  const uint32_t n_train = (uint32_t)o.n_train;
  const uint32_t n_eval = (uint32_t)std::max<uint64_t>(std::min<uint64_t>(o.n_train / 10, 1000000ull), 1000ull);
  const uint32_t total = n_train + n_eval;
  // pair buffers (features, label, candidates) and step buffers: freed at the end
  std::vector<void*> tmp;
  auto talloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
    tmp.push_back(p);
    return p;
  };
  auto cleanup = [&]() {
    for (void* p : tmp) (void)hipFree(p);
  };
  float* feat = static_cast<float*>(talloc((size_t)total * 3 * sizeof(float)));
  int32_t* label = static_cast<int32_t*>(talloc((size_t)total * sizeof(int32_t)));
  uint32_t* t0 = static_cast<uint32_t*>(talloc((size_t)total * sizeof(uint32_t)));
  uint32_t* nc = static_cast<uint32_t*>(talloc((size_t)total * sizeof(uint32_t)));
  uint32_t* count = static_cast<uint32_t*>(talloc(2 * sizeof(uint32_t)));
  StepBuffers sb;
  const bool sb_ok = step_alloc(sb, o.batch, talloc);
  uint32_t* pred = static_cast<uint32_t*>(talloc((size_t)n_eval * sizeof(uint32_t)));
  if (!feat || !label || !t0 || !nc || !count || !sb_ok || !pred) {
    cleanup();
    return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the training set");
  }
  // training pairs first, then held-out pairs from a second stream of Q
  for (int part = 0; part < 2; ++part) {
    const uint32_t cap = part == 0 ? n_train : n_eval;
    const uint64_t off = part == 0 ? 0 : n_train;
    uint32_t got = 0;
    (void)hipMemset(count, 0, sizeof(uint32_t));
    const uint32_t per = 1u << 24;
    for (uint64_t start = 0; got < cap && start < (1ull << 40); start += per) {
      hipLaunchKernelGGL(k_prop_label, dim3(per / 256), dim3(256), 0, 0, G->g, G->prop, o.seed * 2 + part, start, per,
                         cap, count, feat + 3 * off, label + off, t0 + off, nc + off);
      if (hipMemcpy(&got, count, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) {
        cleanup();
        return fail(BRILLE_ERR_DEVICE, "labelling failed");
      }
      if (start == 0 && got == 0) {
        cleanup();
        return fail(BRILLE_ERR_INVALID_ARGUMENT, "no random point landed in a polyhedron node");
      }
    }
    got = std::min(got, cap);
    if (part == 0 && got < cap) {
      cleanup();
      return fail(BRILLE_ERR_DEVICE, "could not draw the training pairs");
    }
  }
  // training
  const uint64_t steps_per_epoch = (n_train + o.batch - 1) / o.batch;
  const uint64_t steps = steps_per_epoch * o.epochs;
  const unsigned nbp = (unsigned)((G->prop_n_params + 255) / 256);
  float last_epoch_loss = 0.f;
  std::vector<float> step_loss(steps_per_epoch);
  for (uint32_t ep = 0; ep < o.epochs; ++ep) {
    (void)hipMemset(sb.loss, 0, sizeof(float));
    for (uint64_t st = 0; st < steps_per_epoch; ++st) {
      const uint32_t b0 = (uint32_t)(st * o.batch);
      const uint32_t n = std::min<uint32_t>(o.batch, n_train - b0);
      grads_launch(G, sb, feat + 3 * (uint64_t)b0, label + b0, t0 + b0, nc + b0, n, 0);
      const uint64_t t = ++G->prop_step;
      const double progress = (double)(ep * steps_per_epoch + st) / (double)steps;
      const float lr = (float)(o.learning_rate * 0.5 * (1.0 + std::cos(3.141592653589793 * progress)));
      const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t)));
      const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t)));
      hipLaunchKernelGGL(k_prop_adam, dim3(nbp), dim3(256), 0, 0, G->prop_params, G->prop_grads, G->prop_m, G->prop_v,
                         G->prop_n_params, lr, c1, c2);
    }
    float sum = 0.f;
    if (hipMemcpy(&sum, sb.loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
      cleanup();
      return fail(BRILLE_ERR_DEVICE, "training failed");
    }
    last_epoch_loss = sum / (float)steps_per_epoch;
  }
  rc = prop_export(G);
  if (rc != BRILLE_OK) {
    cleanup();
    return rc;
  }
  // held-out hit rate with the matrix-core inference
  uint32_t* hits = count + 1;
  (void)hipMemset(hits, 0, sizeof(uint32_t));
  hipLaunchKernelGGL(k_prop_predict, dim3((n_eval + 63) / 64), dim3(256), 0, 0, G->prop, feat + 3 * (uint64_t)n_train,
                     t0 + n_train, nc + n_train, n_eval, pred);
  hipLaunchKernelGGL(k_count_equal, dim3((n_eval + 255) / 256), dim3(256), 0, 0, pred, label + n_train, n_eval,
                     reinterpret_cast<unsigned int*>(hits));
  uint32_t h = 0;
  const hipError_t e = hipMemcpy(&h, hits, sizeof(uint32_t), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(BRILLE_ERR_DEVICE, "evaluation failed");
  if (report) {
    report->n_pairs = n_train;
    report->final_loss = last_epoch_loss;
    report->train_hit_rate = (float)h / (float)n_eval;
    report->n_classes = G->prop_K;
    report->ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  }
  return BRILLE_OK;
}

int brille_proposal_stats(const brille_grid* G, uint64_t* tried, uint64_t* hits) {
  BRILLE_LOCK_GRID(G);
  if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
  if (tried) *tried = G->last_tried;
  if (hits) *hits = G->last_hits;
  return BRILLE_OK;
}

int brille_proposal_loss_grad(brille_grid* G, const float* x, const int32_t* labels, uint32_t n, float* loss,
                              float* grads, uint32_t* n_params) {
  BRILLE_LOCK_GRID(G);
  if (!G || !x || !labels || n == 0) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  if (!G->prop_params) {
    int rc = prop_setup(G, 0);
    if (rc != BRILLE_OK) return rc;
  }
  if (n_params) *n_params = (uint32_t)G->prop_n_params;
  std::vector<uint32_t> t0, nc;
  int rc = host_candidates(G, labels, n, t0, nc);
  if (rc != BRILLE_OK) return rc;
  std::vector<void*> tmp;
  auto talloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
    tmp.push_back(p);
    return p;
  };
  float* dfeat = static_cast<float*>(talloc((size_t)n * 3 * sizeof(float)));
  int32_t* dlab = static_cast<int32_t*>(talloc((size_t)n * sizeof(int32_t)));
  uint32_t* dt0 = static_cast<uint32_t*>(talloc((size_t)n * sizeof(uint32_t)));
  uint32_t* dnc = static_cast<uint32_t*>(talloc((size_t)n * sizeof(uint32_t)));
  StepBuffers sb;
  const bool sb_ok = step_alloc(sb, n, talloc);
  bool ok = dfeat && dlab && dt0 && dnc && sb_ok;
  ok = ok && hipMemcpy(dfeat, x, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(dlab, labels, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(dt0, t0.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(dnc, nc.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemset(sb.loss, 0, sizeof(float)) == hipSuccess;
  if (ok) {
    const unsigned nbp = (unsigned)((G->prop_n_params + 255) / 256);
    hipLaunchKernelGGL(k_fill_zero, dim3(nbp), dim3(256), 0, 0, G->prop_grads, G->prop_n_params);
    grads_launch(G, sb, dfeat, dlab, dt0, dnc, n, 0);
    ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  }
  float l = 0.f;
  ok = ok && hipMemcpy(&l, sb.loss, sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && grads)
    ok = hipMemcpy(grads, G->prop_grads, G->prop_n_params * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  if (ok) {
    const unsigned nbp = (unsigned)((G->prop_n_params + 255) / 256);
    hipLaunchKernelGGL(k_fill_zero, dim3(nbp), dim3(256), 0, 0, G->prop_grads, G->prop_n_params);
    ok = hipDeviceSynchronize() == hipSuccess;
  }
  for (void* p : tmp) (void)hipFree(p);
  if (!ok) return fail(BRILLE_ERR_DEVICE, "proposal loss/gradient failed");
  if (loss) *loss = l;
  return BRILLE_OK;
}

int brille_proposal_get_params(const brille_grid* G, float* params, uint32_t* n_params) {
  BRILLE_LOCK_GRID(G);
  if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
  if (!G->prop_params) return fail(BRILLE_ERR_INVALID_ARGUMENT, "the proposal has no parameters yet");
  if (n_params) *n_params = (uint32_t)G->prop_n_params;
  if (params && hipMemcpy(params, G->prop_params, G->prop_n_params * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpy failed");
  return BRILLE_OK;
}

int brille_proposal_set_params(brille_grid* G, const float* params, uint32_t n_params) {
  BRILLE_LOCK_GRID(G);
  if (!G || !params) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  if (!G->prop_params) {
    int rc = prop_setup(G, 0);
    if (rc != BRILLE_OK) return rc;
  }
  if (n_params != G->prop_n_params)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "expected " + std::to_string(G->prop_n_params) + " parameters");
  if (hipMemcpy(G->prop_params, params, G->prop_n_params * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipMemcpy failed");
  return prop_export(G);
}

int brille_proposal_predict(brille_grid* G, const float* x, const int32_t* in_tet, uint32_t n, int32_t* out) {
  BRILLE_LOCK_GRID(G);
  if (!G || !x || !in_tet || !out) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  if (!G->prop_trained) return fail(BRILLE_ERR_INVALID_ARGUMENT, "the proposal has not been trained or set");
  if (n == 0) return BRILLE_OK;
  std::vector<uint32_t> t0, nc;
  int rc = host_candidates(G, in_tet, n, t0, nc);
  if (rc != BRILLE_OK) return rc;
  float* dfeat = nullptr;
  uint32_t *dt0 = nullptr, *dnc = nullptr, *dout = nullptr;
  bool ok = hipMalloc(&dfeat, (size_t)n * 3 * sizeof(float)) == hipSuccess &&
            hipMalloc(&dt0, (size_t)n * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&dnc, (size_t)n * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&dout, (size_t)n * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMemcpy(dfeat, x, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(dt0, t0.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(dnc, nc.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_prop_predict, dim3((n + 63) / 64), dim3(256), 0, 0, G->prop, dfeat, dt0, dnc, n, dout);
    ok = hipGetLastError() == hipSuccess;
  }
  ok = ok && hipMemcpy(out, dout, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess;
  for (void* p : {(void*)dfeat, (void*)dt0, (void*)dnc, (void*)dout})
    if (p) (void)hipFree(p);
  if (!ok) return fail(BRILLE_ERR_DEVICE, "proposal inference failed");
  return BRILLE_OK;
}

}  // extern "C"
