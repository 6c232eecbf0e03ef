// Learned tetrahedron proposal: inference on the matrix cores.
//
// The network (3 -> 128 -> 128 -> K, ReLU) proposes which tetrahedron of a
// polyhedron node holds a point.  The proposal is only a hint: the tile
// kernels verify it with the reference's own FP64 predicates and run the
// exact scan of PolyNode::indices_weights (src/trellis_node.hpp:275-308) when
// it does not pass, so the chosen tetrahedron, the indices and the weights are
// always the reference's.
//
// One wave evaluates 16 points at a time:
//   layer 1 (3 -> 128, FP32) is computed straight into the A operands of
//     v_mfma_f32_16x16x32_bf16: lane l holds H1[i = l%16][k = 32 kb + 8 (l/16) + e];
//   layer 2 (128 -> 128) is 8 column blocks x 4 k-blocks of that MFMA, with
//     the bf16 rows of W2 as B operands (B[k][j] = W2[j][k], lane l holds
//     k = 32 kb + 8 (l/16) + e, j = 16 nb + l%16); D[i][j] lands in lane l as
//     i = 4 (l/16) + r, j = 16 nb + l%16, r = 0..3;
//   layer 3 is evaluated only for the candidate tetrahedra of each point: the
//     16 lanes of a group hold one point's 128 hidden units (8 per lane) and
//     reduce W3[c] . h2 with four xor-shuffles.  W3 rows are stored
//     lane-major so that each lane reads its 8 weights with one 16-byte load.
#pragma once
#include "exact_path.hpp"

namespace brille_amd {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

// network input of a point: its Cartesian coordinates inside its trellis
// node, scaled into [-1, 1] (a node's tetrahedra are a fan from one apex, so
// their boundaries are a few planes in these coordinates)
__device__ __forceinline__ void prop_features(const GridConst& g, const double* x, uint32_t node, float* f) {
  const uint32_t s0 = g.nk[0] - 1, s1 = g.nk[1] - 1;
  const uint32_t sub[3] = {node % s0, (node / s0) % s1, node / (s0 * s1)};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double lo = g.knots[d][sub[d]], hi = g.knots[d][sub[d] + 1];
    f[d] = (float)(2.0 * (x[d] - lo) / (hi - lo) - 1.0);
  }
}

// Layers 1 and 2 read W1, b1, W2 and b2 through W (global memory, or LDS
// copies held by the block, which keep them out of the vector memory path).
template <typename F, typename H>
struct PropWeights {
  F W1, b1, b2;
  H W2;
};
using LdsF = const __attribute__((address_space(3))) float*;
using LdsH = const __attribute__((address_space(3))) uint16_t*;
__device__ __forceinline__ bf16x8_t load_bf16x8(const uint16_t* p) { return *reinterpret_cast<const bf16x8_t*>(p); }
__device__ __forceinline__ bf16x8_t load_bf16x8(LdsH p) {
  return *reinterpret_cast<const __attribute__((address_space(3))) bf16x8_t*>(p);
}

// Proposals for the n <= 16 points of a chunk (feat: 3 floats per point,
// t0/nc: first candidate and candidate count).  Called by a whole wave with
// wave-uniform arguments; out[i] receives the first candidate of largest logit.
template <typename F, typename H>
__device__ __forceinline__ void propose16_inline(const ProposalConst& pc, PropWeights<F, H> W, const float* feat,
                                                 const uint32_t* t0, const uint32_t* nc, uint32_t n, uint32_t* out) {
  const uint32_t l = threadIdx.x & 63u, m = l & 15u, grp = l >> 4;
  float x0 = 0.f, x1 = 0.f, x2 = 0.f;
  if (m < n) {
    x0 = feat[3 * m];
    x1 = feat[3 * m + 1];
    x2 = feat[3 * m + 2];
  }
  bf16x8_t a[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = kb * 32 + 8 * (int)grp + e;
      float h = W.W1[3 * k] * x0;
      h += W.W1[3 * k + 1] * x1;
      h += W.W1[3 * k + 2] * x2;
      h += W.b1[k];
      a[kb][e] = (__bf16)(h > 0.f ? h : 0.f);
    }
  floatx4_t acc[8];
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    acc[nb] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    const uint32_t w0 = (nb * 16 + m) * kPropH + 8 * grp;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kb], load_bf16x8(W.W2 + w0 + 32 * kb), acc[nb], 0, 0, 0);
    const float bj = W.b2[nb * 16 + m];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = acc[nb][r] + bj;
      acc[nb][r] = v > 0.f ? v : 0.f;
    }
  }
  uint32_t ncr[4], t0r[4], bestc[4];
  float best[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t pi = 4 * grp + r;
    ncr[r] = pi < n ? nc[pi] : 0u;
    t0r[r] = pi < n ? t0[pi] : 0u;
    bestc[r] = t0r[r];
    best[r] = -3.0e38f;
  }
  uint32_t maxc = 0;
  for (uint32_t pi = 0; pi < n; ++pi) maxc = nc[pi] > maxc ? nc[pi] : maxc;
  // candidates two at a time: the 8 row loads of a step are issued together
  constexpr int CQ = 2;
  for (uint32_t s0 = 0; s0 < maxc; s0 += CQ) {
    bf16x8_t w[CQ][4];
    float bias[CQ][4];
#pragma unroll
    for (int q = 0; q < CQ; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t s = s0 + q;
        const uint32_t c = s < ncr[r] ? t0r[r] + s : t0r[r];  // a valid row either way
        w[q][r] = *reinterpret_cast<const bf16x8_t*>(pc.W3 + (size_t)c * kPropH + 8 * m);
        bias[q][r] = pc.b3[c];
      }
#pragma unroll
    for (int q = 0; q < CQ; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t s = s0 + q;
        float part = 0.f;
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) part += (float)w[q][r][nb] * acc[nb][r];
        part += __shfl_xor(part, 1);
        part += __shfl_xor(part, 2);
        part += __shfl_xor(part, 4);
        part += __shfl_xor(part, 8);
        const float z = part + bias[q][r];
        if (s < ncr[r] && z > best[r]) {
          best[r] = z;
          bestc[r] = t0r[r] + s;
        }
      }
  }
  if (m == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t pi = 4 * grp + r;
      if (pi < n) out[pi] = bestc[r];
    }
  }
}

// Out-of-line form for callers that loop over chunks (inlined into such a
// loop, the per-lane layer-1 weights would be hoisted and held across
// iterations, over a hundred registers).
template <typename F, typename H>
__device__ __noinline__ void propose16(const ProposalConst& pc, PropWeights<F, H> W, const float* feat,
                                       const uint32_t* t0, const uint32_t* nc, uint32_t n, uint32_t* out) {
  propose16_inline(pc, W, feat, t0, nc, n, out);
}

}  // namespace brille_amd
