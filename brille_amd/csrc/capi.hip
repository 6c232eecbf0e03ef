// C ABI of the MI355X ir_interpolate_at path (declarations in include/brille_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/brille_amd.h"
#include "grid_state.hpp"
#include "interp_kernels.hpp"

using namespace brille_amd;

namespace brille_amd {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

const char* g_last_error_text() { return g_last_error.c_str(); }

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(BRILLE_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int upload(const T* host, size_t n, const T** dev, std::vector<void*>& allocs) {
  if (n == 0 || host == nullptr) {
    *dev = nullptr;
    return BRILLE_OK;
  }
  void* p = nullptr;
  HIP_TRY(hipMalloc(&p, n * sizeof(T)));
  allocs.push_back(p);
  HIP_TRY(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
  *dev = static_cast<const T*>(p);
  return BRILLE_OK;
}

static void free_all(std::vector<void*>& allocs) {
  for (void* p : allocs) (void)hipFree(p);
  allocs.clear();
}

// reals per payload row: 16-byte multiple, and a power of two up to a cache
// line so that no row of 128 bytes or less straddles two lines
static uint32_t row_stride(uint32_t used, size_t real_bytes) {
  size_t bytes = ((size_t)used * real_bytes + 15) / 16 * 16;
  if (bytes <= 128) {
    size_t p = 16;
    while (p < bytes) p <<= 1;
    bytes = p;
  }
  return (uint32_t)(bytes / real_bytes);
}

// Packing of payload rows: row r = (vertex, branch) holds the vectors'
// complex elements (re, im interleaved) then the values, zero padded to the
// stride.  One thread per output real; In -> Out converts by rounding to
// nearest (a static_cast).
constexpr uint64_t kFillPiece = 256ull << 20;  // bytes of caller data staged per packing launch

template <typename In, typename Out>
__global__ __launch_bounds__(256) void k_pack_rows(const In* __restrict__ vals, const In* __restrict__ vecs,
                                                   uint64_t m, uint32_t sv, uint32_t sc, uint32_t stride,
                                                   Out* __restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * stride) return;
  const uint64_t r = t / stride;
  const uint32_t k = (uint32_t)(t - r * stride);
  In x = (In)0;
  if (k < 2 * sc) x = vecs[r * 2 * sc + k];
  else if (k < 2 * sc + sv) x = vals[r * sv + (k - 2 * sc)];
  out[t] = static_cast<Out>(x);
}

template <typename In>
static void pack_both(const In* v, const In* c, uint64_t m, const DataConst& d, double* o64, float* o32) {
  if (o64) {
    const uint64_t n = m * d.stride64;
    hipLaunchKernelGGL((k_pack_rows<In, double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, v, c, m, d.sv,
                       d.sc, d.stride64, o64);
  }
  if (o32) {
    const uint64_t n = m * d.stride32;
    hipLaunchKernelGGL((k_pack_rows<In, float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, v, c, m, d.sv,
                       d.sc, d.stride32, o32);
  }
}

// rows [r0, r0 + m) from values v and vectors c (device memory, data_type reals)
static int pack_rows(const DataConst& d, const void* v, const void* c, int data_type, uint64_t r0, uint64_t m,
                     double* rows64, float* rows32) {
  double* o64 = rows64 ? rows64 + r0 * d.stride64 : nullptr;
  float* o32 = rows32 ? rows32 + r0 * d.stride32 : nullptr;
  if (data_type == BRILLE_FP32) pack_both(static_cast<const float*>(v), static_cast<const float*>(c), m, d, o64, o32);
  else pack_both(static_cast<const double*>(v), static_cast<const double*>(c), m, d, o64, o32);
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "payload packing launch failed");
  return BRILLE_OK;
}

static void make_tols(double Ttol, int tol, double* rel, double* abs_) {
  // approx_float::tols  src/approx_float.hpp:79-100
  double Trel = std::numeric_limits<double>::epsilon();
  double Tabs = 5.0 / 1000000000000000.0;
  Trel *= static_cast<double>(tol) * 10000.0;
  if (Ttol > Trel) Trel = Ttol;
  if (Ttol > Tabs) Tabs = Ttol;
  *rel = Trel;
  *abs_ = Tabs;
}

__global__ void k_err_init(ErrRec* e) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    e->first_bad = ~0ull;
    e->count = 0ull;
  }
}

// stage (a) only
__global__ __launch_bounds__(256) void k_ir_moveinto(GridConst g, const double* Q, uint64_t n, double* q_out,
                                                     int32_t* tau_out, uint32_t* rot_out, uint32_t* invrot_out,
                                                     uint8_t* status_out, ErrRec* err, uint64_t base) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const Tol tbz{g.tol_bz_rel, g.tol_bz_abs};
  const Tol tdef{g.tol_def_rel, g.tol_def_abs};
  double Qp[3] = {Q[3 * p], Q[3 * p + 1], Q[3 * p + 2]};
  double q[3] = {0, 0, 0};
  int32_t tau[3] = {0, 0, 0};
  uint32_t r = g.identity, ir = g.identity;
  int st = moveinto_exact(g, Qp, q, tau, tbz);
  if (st == BRILLE_PT_OK) st = wedge_search_exact(g, q, &r, &ir, tbz, tdef);
  if (q_out) { q_out[3 * p] = q[0]; q_out[3 * p + 1] = q[1]; q_out[3 * p + 2] = q[2]; }
  if (tau_out) { tau_out[3 * p] = tau[0]; tau_out[3 * p + 1] = tau[1]; tau_out[3 * p + 2] = tau[2]; }
  if (rot_out) rot_out[p] = r;
  if (invrot_out) invrot_out[p] = ir;
  if (status_out) status_out[p] = (uint8_t)st;
  if (st != BRILLE_PT_OK) {
    atomicMin(&err->first_bad, err_key(base + p, st));
    atomicAdd(&err->count, 1ull);
  }
}

}  // namespace brille_amd

// ===========================================================================
extern "C" {

int brille_abi_version(void) { return BRILLE_AMD_ABI_VERSION; }

int brille_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

size_t brille_last_error(char* buffer, size_t length) {
  const std::string& m = g_last_error;
  if (buffer && length) {
    size_t k = std::min(length - 1, m.size());
    std::memcpy(buffer, m.data(), k);
    buffer[k] = '\0';
  }
  return m.size();
}

// Every index the kernels follow is checked here, once, so that no table can
// send a kernel outside an allocation.
static int validate_tables(const brille_grid_desc* d) {
  auto bad = [](const std::string& m) { return fail(BRILLE_ERR_INVALID_ARGUMENT, m); };
  for (uint32_t j = 0; j < d->n_ops; ++j)
    if (d->inverse_index[j] >= d->n_ops) return bad("inverse_index out of range");
  for (uint32_t j = 0; j < d->n_faces; ++j)
    if (!(d->prim_tau_lens[j] > 0)) return bad("first-zone tau lengths must be positive");
  if (d->n_nodes == 0) return BRILLE_OK;
  for (int a = 0; a < 3; ++a) {
    if (d->n_knots[a] < 2) return bad("each trellis axis needs at least two knots");
    for (uint32_t k = 1; k < d->n_knots[a]; ++k)
      if (!(d->knots[a][k] > d->knots[a][k - 1])) return bad("trellis knots must increase strictly");
  }
  for (uint64_t i = 0; i < d->n_nodes; ++i) {
    const uint8_t t = d->node_type[i];
    if (t > BRILLE_NODE_POLY) return bad("unknown node type");
    if (t == BRILLE_NODE_CUBE && d->node_index[i] >= d->n_cubes) return bad("cube node index out of range");
    if (t == BRILLE_NODE_POLY && d->node_index[i] >= d->n_polys) return bad("poly node index out of range");
  }
  for (uint64_t i = 0; i < 8 * d->n_cubes; ++i)
    if (d->cube_vertices[i] >= d->n_vertices) return bad("cube vertex index out of range");
  if (d->n_polys) {
    if (d->poly_tet_offset[0] != 0 || d->poly_tet_offset[d->n_polys] != d->n_tets) return bad("poly offsets do not cover the tetrahedra");
    for (uint64_t p = 0; p < d->n_polys; ++p)
      if (d->poly_tet_offset[p + 1] < d->poly_tet_offset[p]) return bad("poly offsets must not decrease");
  }
  for (uint64_t i = 0; i < 4 * d->n_tets; ++i)
    if (d->tet_vertices[i] >= d->n_vertices) return bad("tetrahedron vertex index out of range");
  for (uint64_t i = 0; i < d->n_tets; ++i)
    if (!(d->tet_volume[i] > 0)) return bad("tetrahedron volumes must be positive");
  if (d->n_tets >= 0xFFFFFFFFull || d->n_vertices >= 0xFFFFFFFFull) return bad("tables exceed 32-bit indices");
  return BRILLE_OK;
}

int brille_grid_create(const brille_grid_desc* desc, brille_grid** out) {
  if (!desc || !out) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  if (desc->n_faces < 4 || desc->n_faces > kMaxFaces)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "first Brillouin zone must have between 4 and 32 faces");
  if (desc->n_ops < 1 || desc->identity_index >= desc->n_ops)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "point group must contain the identity");
  if (int v = validate_tables(desc); v != BRILLE_OK) return v;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(BRILLE_ERR_NO_DEVICE, "no HIP device available");
  auto* G = new brille_grid();
  GridConst& g = G->g;
  std::memset(&g, 0, sizeof(g));
  std::memcpy(g.outer_to_xyz, desc->outer_to_xyz, sizeof(g.outer_to_xyz));
  std::memcpy(g.outer_Gr, desc->outer_metric_recip, sizeof(g.outer_Gr));
  std::memcpy(g.outer_Gd, desc->outer_metric_real, sizeof(g.outer_Gd));
  g.outer_scale = desc->outer_cross_scale;
  std::memcpy(g.prim_Gr, desc->prim_metric_recip, sizeof(g.prim_Gr));
  std::memcpy(g.prim_Gd, desc->prim_metric_real, sizeof(g.prim_Gd));
  g.prim_scale = desc->prim_cross_scale;
  std::memcpy(g.P6t, desc->to_primitive_6Pt, sizeof(g.P6t));
  std::memcpy(g.invPt, desc->from_primitive_invPt, sizeof(g.invPt));
  g.transform_needed = desc->transform_needed;
  g.no_ir_mirroring = desc->no_ir_mirroring;
  g.has_seed = desc->has_wedge_seed;
  g.F = desc->n_faces;
  g.W = desc->n_wedge;
  g.G = desc->n_ops;
  g.identity = desc->identity_index;
  make_tols(desc->float_tolerance, desc->approx_digits, &g.tol_bz_rel, &g.tol_bz_abs);
  make_tols(0.0, 1, &g.tol_def_rel, &g.tol_def_abs);
  auto& A = G->grid_allocs;
  const size_t F3 = 3 * (size_t)desc->n_faces;
  int rc = BRILLE_OK;
  auto chk = [&](int r) { if (r != BRILLE_OK && rc == BRILLE_OK) rc = r; };
  chk(upload(desc->prim_plane_a, F3, &g.prim_pa, A));
  chk(upload(desc->prim_plane_b, F3, &g.prim_pb, A));
  chk(upload(desc->prim_plane_c, F3, &g.prim_pc, A));
  chk(upload(desc->prim_normals, F3, &g.prim_n, A));
  chk(upload(desc->prim_taus, F3, &g.prim_taus, A));
  chk(upload(desc->prim_tau_lens, (size_t)desc->n_faces, &g.prim_tau_lens, A));
  chk(upload(desc->outer_plane_a, F3, &g.outer_pa, A));
  chk(upload(desc->outer_plane_b, F3, &g.outer_pb, A));
  chk(upload(desc->outer_plane_c, F3, &g.outer_pc, A));
  chk(upload(desc->wedge_normals, 3 * (size_t)desc->n_wedge, &g.wedge_n, A));
  chk(upload(desc->rotations, 9 * (size_t)desc->n_ops, &g.rot, A));
  chk(upload(desc->inverse_index, (size_t)desc->n_ops, &g.inv_index, A));
  // trellis (may be empty for a zone-only handle)
  uint64_t nodes = 1;
  for (int d = 0; d < 3; ++d) {
    g.nk[d] = desc->n_knots[d];
    if (desc->n_knots[d] >= 2) {
      chk(upload(desc->knots[d], (size_t)desc->n_knots[d], &g.knots[d], A));
      g.knot_lo[d] = desc->knots[d][0];
      double step = (desc->knots[d][desc->n_knots[d] - 1] - desc->knots[d][0]) / (double)(desc->n_knots[d] - 1);
      g.knot_inv_step[d] = step > 0 ? 1.0 / step : 0.0;
      nodes *= desc->n_knots[d] - 1;
    } else {
      nodes = 0;
    }
  }
  if (desc->n_nodes && nodes != desc->n_nodes) {
    free_all(A);
    delete G;
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "node count does not match the knots");
  }
  g.n_nodes = desc->n_nodes;
  chk(upload(desc->node_type, (size_t)desc->n_nodes, &g.node_type, A));
  chk(upload(desc->node_index, (size_t)desc->n_nodes, &g.node_index, A));
  chk(upload(desc->cube_vertices, 8 * (size_t)desc->n_cubes, &g.cube_v, A));
  if (desc->n_polys) chk(upload(desc->poly_tet_offset, (size_t)desc->n_polys + 1, &g.poly_off, A));
  chk(upload(desc->tet_vertices, 4 * (size_t)desc->n_tets, &g.tet_v, A));
  chk(upload(desc->tet_circumsphere, 4 * (size_t)desc->n_tets, &g.tet_c, A));
  chk(upload(desc->tet_volume, (size_t)desc->n_tets, &g.tet_vol, A));
  chk(upload(desc->vertices, 3 * (size_t)desc->n_vertices, &g.verts, A));
  if (rc == BRILLE_OK) rc = build_fast_tables(desc, G);
  if (rc != BRILLE_OK) {
    free_all(A);
    delete G;
    return rc;
  }
  G->n_vertices = desc->n_vertices;
  G->n_cubes = desc->n_cubes;
  G->n_polys = desc->n_polys;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      G->n_cu = (uint32_t)prop.multiProcessorCount;
  }
  G->n_tets = desc->n_tets;
  void* e = nullptr;
  if (hipMalloc(&e, sizeof(ErrRec)) != hipSuccess) {
    free_all(A);
    delete G;
    return fail(BRILLE_ERR_DEVICE, "hipMalloc failed for the error record");
  }
  G->err = static_cast<ErrRec*>(e);
  G->host_desc_poly_off.assign(desc->poly_tet_offset, desc->poly_tet_offset + (desc->n_polys ? desc->n_polys + 1 : 0));
  *out = G;
  return BRILLE_OK;
}

int brille_grid_destroy(brille_grid* G) {
  if (!G) return BRILLE_OK;
  { std::lock_guard<std::mutex> wait_for_calls(G->mu); }
  (void)hipDeviceSynchronize();
  free_all(G->grid_allocs);
  free_all(G->data_allocs);
  free_all(G->sort_allocs);
  free_proposal(G);
  free_staging(G);
  free_binned(G);
  if (G->err) (void)hipFree(G->err);
  delete G;
  return BRILLE_OK;
}

// The payload is packed into rows on the device: the caller's values (V, B,
// sv) and complex vectors (V, B, sc), host or device, FP64 or FP32, go
// through packing kernels in pieces of at most kFillPiece bytes of input, so
// no host copy of the rows is ever made (the largest payloads are tens of GB).
static int fill_payload(brille_grid* G, const brille_data_desc* vals, const brille_data_desc* vecs,
                        const brille_fill_opts& o) {
  DataConst& d = G->d;
  auto& A = G->data_allocs;
  const uint64_t R = (uint64_t)d.V * d.B;
  const size_t in_real = o.data_type == BRILLE_FP32 ? 4 : 8;
  const bool keep64 = o.keep == 0 || (o.keep & BRILLE_KEEP_FP64);
  const bool keep32 = o.keep == 0 || (o.keep & BRILLE_KEEP_FP32);
  void* out64 = nullptr;
  void* out32 = nullptr;
  if (keep64) {
    HIP_TRY(hipMalloc(&out64, R * d.stride64 * sizeof(double) + 16));
    A.push_back(out64);
    d.rows64 = static_cast<const double*>(out64);
  }
  if (keep32) {
    HIP_TRY(hipMalloc(&out32, R * d.stride32 * sizeof(float) + 16));
    A.push_back(out32);
    d.rows32 = static_cast<const float*>(out32);
  }
  if (R == 0) return BRILLE_OK;
  const size_t row_in = (size_t)(2 * d.sc + d.sv) * in_real;
  const uint64_t piece = std::max<uint64_t>(1, kFillPiece / row_in);
  const char* hv = reinterpret_cast<const char*>(vals->data);
  const char* hc = reinterpret_cast<const char*>(vecs->data);
  void* stage = nullptr;
  if (o.memory == BRILLE_MEM_HOST) HIP_TRY(hipMalloc(&stage, std::min<uint64_t>(piece, R) * row_in + 256));
  int rc = BRILLE_OK;
  for (uint64_t r0 = 0; r0 < R && rc == BRILLE_OK; r0 += piece) {
    const uint64_t m = std::min(piece, R - r0);
    const size_t vb = (size_t)m * d.sv * in_real, cb = (size_t)m * 2 * d.sc * in_real;
    const void* pv = hv + r0 * d.sv * in_real;
    const void* pc = hc + r0 * 2 * d.sc * in_real;
    if (o.memory == BRILLE_MEM_HOST) {
      char* st = static_cast<char*>(stage);
      if (hipMemcpy(st, pc, cb, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(st + ((cb + 255) & ~size_t(255)), pv, vb, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(BRILLE_ERR_DEVICE, "hipMemcpy failed while filling the payload");
        break;
      }
      pc = st;
      pv = st + ((cb + 255) & ~size_t(255));
    }
    rc = pack_rows(d, pv, pc, o.data_type, r0, m, static_cast<double*>(out64), static_cast<float*>(out32));
  }
  if (stage) (void)hipFree(stage);
  if (rc == BRILLE_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(BRILLE_ERR_DEVICE, "payload packing failed");
  return rc;
}

static int grid_fill(brille_grid* G, const brille_data_desc* vals, const brille_data_desc* vecs,
                     const brille_gamma_desc* gamma, const brille_perm_desc* perm, const brille_fill_opts& o) {
  // every check first: a fill that fails them leaves the previous payload in place
  if (!G || !vals || !vecs) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  if (vals->n_vertices != G->n_vertices || vecs->n_vertices != G->n_vertices)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "Provided " + std::to_string(vals->n_vertices) + " arrays but " +
                                                  std::to_string(G->n_vertices) + " were expected!");
  if (vals->branches != vecs->branches || vals->branches == 0)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "Inconsistent values and vectors provided to DualInterpolator");
  if (vals->is_complex || !vecs->is_complex)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "values must be real and vectors complex");
  if (o.memory != BRILLE_MEM_HOST && o.memory != BRILLE_MEM_DEVICE)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "unknown payload memory");
  if (o.data_type != BRILLE_FP64 && o.data_type != BRILLE_FP32)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "unknown payload data type");
  if (o.keep & ~(BRILLE_KEEP_FP64 | BRILLE_KEEP_FP32))
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "unknown payload precision to keep");
  for (const brille_data_desc* x : {vals, vecs}) {
    if (x->elements[1] % 3) return fail(BRILLE_ERR_INVALID_ARGUMENT, "Vectors must have 3N elements per branch");
    if (x->elements[2] % 9) return fail(BRILLE_ERR_INVALID_ARGUMENT, "Matrices must have 9N elements per branch");
    if (x->elements[0] + x->elements[1] + x->elements[2] == 0)
      return fail(BRILLE_ERR_INVALID_ARGUMENT, "the elements per branch must not all be zero");
    if (G->n_vertices && !x->data) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null payload data");
  }
  if (gamma) {
    // the kernels store to output atom 3 F0 + i and read vectors[vector_index]:
    // every entry is checked before anything is uploaded
    if (gamma->n_ops != G->g.G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "GammaTable and point group sizes differ");
    const uint64_t m = (uint64_t)gamma->n_atoms * gamma->n_ops;
    if (m && (!gamma->F0 || !gamma->vector_index))
      return fail(BRILLE_ERR_INVALID_ARGUMENT, "GammaTable mappings must not be null");
    if (gamma->n_vectors && !gamma->vectors)
      return fail(BRILLE_ERR_INVALID_ARGUMENT, "GammaTable vectors must not be null");
    for (uint64_t i = 0; i < m; ++i) {
      if (gamma->F0[i] >= gamma->n_atoms)
        return fail(BRILLE_ERR_INVALID_ARGUMENT, "GammaTable atom mapping out of range");
      if (gamma->vector_index[i] >= gamma->n_vectors)
        return fail(BRILLE_ERR_INVALID_ARGUMENT, "GammaTable vector index out of range");
    }
  }
  if (perm && perm->n_keys) {
    if (!perm->keys || !perm->perm_index || (perm->n_perms && !perm->perms))
      return fail(BRILLE_ERR_INVALID_ARGUMENT, "permutation table arrays must not be null");
    for (uint64_t i = 0; i < perm->n_keys; ++i) {
      if (perm->perm_index[i] >= perm->n_perms) return fail(BRILLE_ERR_INVALID_ARGUMENT, "permutation index out of range");
      if (i && perm->keys[i] <= perm->keys[i - 1]) return fail(BRILLE_ERR_INVALID_ARGUMENT, "permutation keys must be sorted");
    }
    for (uint64_t i = 0; i < (uint64_t)perm->n_perms * vals->branches; ++i)
      if (perm->perms[i] >= vals->branches) return fail(BRILLE_ERR_INVALID_ARGUMENT, "permutation entry out of range");
  }
  (void)hipDeviceSynchronize();  // nothing still reads the payload that is replaced
  free_all(G->data_allocs);
  free_all(G->sort_allocs);
  G->n_perms = 0;
  ++G->fill_gen;  // the payload tables are replaced from here on (batches re-read them)
  G->filled = false;
  DataConst& d = G->d;
  std::memset(&d, 0, sizeof(d));
  d.V = G->n_vertices;
  d.B = vals->branches;
  d.sv = vals->elements[0] + vals->elements[1] + vals->elements[2];
  d.sc = vecs->elements[0] + vecs->elements[1] + vecs->elements[2];
  for (int i = 0; i < 3; ++i) { d.ev[i] = vals->elements[i]; d.ec[i] = vecs->elements[i]; }
  d.val_rot = vals->rotates_like;
  d.val_lu = vals->length_unit;
  d.vec_rot = vecs->rotates_like;
  d.vec_lu = vecs->length_unit;
  auto& A = G->data_allocs;
  const uint32_t used = 2 * d.sc + d.sv;
  d.stride64 = row_stride(used, sizeof(double));
  d.stride32 = row_stride(used, sizeof(float));
  int rc = fill_payload(G, vals, vecs, o);
  auto chk = [&](int r) { if (r != BRILLE_OK && rc == BRILLE_OK) rc = r; };
  if (gamma) {
    d.n_atoms = gamma->n_atoms;
    d.gamma_ops = gamma->n_ops;
    chk(upload(gamma->F0, (size_t)gamma->n_atoms * gamma->n_ops, &d.F0, A));
    chk(upload(gamma->vector_index, (size_t)gamma->n_atoms * gamma->n_ops, &d.vidx, A));
    chk(upload(gamma->vectors, 3 * (size_t)gamma->n_vectors, &d.gvec, A));
    std::vector<double> gkey(3 * (size_t)gamma->n_atoms * gamma->n_ops);
    for (size_t key = 0; key < (size_t)gamma->n_atoms * gamma->n_ops; ++key)
      for (int c = 0; c < 3; ++c) gkey[3 * key + c] = gamma->vectors[3 * (size_t)gamma->vector_index[key] + c];
    chk(upload(gkey.data(), gkey.size(), &d.gkey, A));
    if (gamma->cart_rotations) chk(upload(gamma->cart_rotations, 9 * (size_t)gamma->n_ops, &d.cart_rot, A));
  }
  if (perm && perm->n_keys) {
    d.n_keys = perm->n_keys;
    chk(upload(perm->keys, (size_t)perm->n_keys, &d.keys, A));
    chk(upload(perm->perm_index, (size_t)perm->n_keys, &d.perm_index, A));
    chk(upload(perm->perms, (size_t)perm->n_perms * d.B, &d.perms, A));
    G->n_perms = perm->n_perms;
  }
  if (rc != BRILLE_OK) {
    free_all(A);
    std::memset(&d, 0, sizeof(d));
    G->filled = false;
    return rc;
  }
  G->filled = true;
  return BRILLE_OK;
}

int brille_grid_fill(brille_grid* G, const brille_data_desc* vals, const brille_data_desc* vecs,
                     const brille_gamma_desc* gamma, const brille_perm_desc* perm) {
  BRILLE_LOCK_GRID(G);
  brille_fill_opts o{};
  o.memory = BRILLE_MEM_HOST;
  o.data_type = BRILLE_FP64;
  o.keep = 0;
  return grid_fill(G, vals, vecs, gamma, perm, o);
}

int brille_grid_fill_ex(brille_grid* G, const brille_data_desc* vals, const brille_data_desc* vecs,
                        const brille_gamma_desc* gamma, const brille_perm_desc* perm, const brille_fill_opts* opts) {
  BRILLE_LOCK_GRID(G);
  if (!opts) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null fill options");
  return grid_fill(G, vals, vecs, gamma, perm, *opts);
}

uint64_t brille_bytes_per_point(const brille_grid* G, int32_t precision) {
  BRILLE_LOCK_GRID(G);
  if (!G || !G->filled) return 0;
  const uint64_t s = precision == BRILLE_FP32 ? 4 : 8;
  return (uint64_t)G->d.B * G->d.sv * s + (uint64_t)G->d.B * G->d.sc * 2 * s;
}

int brille_ir_interpolate_at(brille_grid* G, const double* Q, uint64_t n, const brille_run_opts* opts_in,
                             void* values_out, void* vectors_out, const brille_aux_out* aux, void* stream) {
  BRILLE_LOCK_GRID(G);
  return run_interpolate(G, Q, n, opts_in, values_out, vectors_out, aux, static_cast<hipStream_t>(stream), nullptr, 0);
}

int brille_interpolate_at(brille_grid* G, const double* Q, uint64_t n, const brille_run_opts* opts_in,
                          void* values_out, void* vectors_out, const brille_aux_out* aux, void* stream) {
  BRILLE_LOCK_GRID(G);
  return run_interpolate(G, Q, n, opts_in, values_out, vectors_out, aux, static_cast<hipStream_t>(stream), nullptr, 0,
                         FOLD_FIRST_ZONE);
}

int brille_ir_moveinto(brille_grid* G, const double* Q, uint64_t n, int32_t memory, double* q_out, int32_t* tau_out,
                       uint32_t* rot_out, uint32_t* invrot_out, uint8_t* status_out, void* stream_v) {
  BRILLE_LOCK_GRID(G);
  if (!G || (!Q && n)) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return BRILLE_OK;
  hipStream_t s = static_cast<hipStream_t>(stream_v);
  const double* dQ = Q;
  double* dq = q_out;
  int32_t* dtau = tau_out;
  uint32_t *drot = rot_out, *dinv = invrot_out;
  uint8_t* dst = status_out;
  std::vector<void*> tmp;
  auto dalloc = [&](void** p, size_t bytes) -> int {
    HIP_TRY(hipMalloc(p, bytes));
    tmp.push_back(*p);
    return BRILLE_OK;
  };
  int rc = BRILLE_OK;
  if (memory == BRILLE_MEM_HOST) {
    void* p;
    if ((rc = dalloc(&p, n * 3 * sizeof(double))) != BRILLE_OK) { free_all(tmp); return rc; }
    HIP_TRY(hipMemcpy(p, Q, n * 3 * sizeof(double), hipMemcpyHostToDevice));
    dQ = static_cast<const double*>(p);
    if (q_out) { if ((rc = dalloc(&p, n * 3 * sizeof(double))) != BRILLE_OK) { free_all(tmp); return rc; } dq = (double*)p; }
    if (tau_out) { if ((rc = dalloc(&p, n * 3 * sizeof(int32_t))) != BRILLE_OK) { free_all(tmp); return rc; } dtau = (int32_t*)p; }
    if (rot_out) { if ((rc = dalloc(&p, n * sizeof(uint32_t))) != BRILLE_OK) { free_all(tmp); return rc; } drot = (uint32_t*)p; }
    if (invrot_out) { if ((rc = dalloc(&p, n * sizeof(uint32_t))) != BRILLE_OK) { free_all(tmp); return rc; } dinv = (uint32_t*)p; }
    if (status_out) { if ((rc = dalloc(&p, n)) != BRILLE_OK) { free_all(tmp); return rc; } dst = (uint8_t*)p; }
  }
  hipLaunchKernelGGL(k_err_init, dim3(1), dim3(1), 0, s, G->err);
  const uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_ir_moveinto, dim3((unsigned)blocks), dim3(256), 0, s, G->g, dQ, n, dq, dtau, drot, dinv, dst, G->err,
                     (uint64_t)0);
  HIP_TRY(hipGetLastError());
  ErrRec e;
  HIP_TRY(hipMemcpyAsync(&e, G->err, sizeof(e), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (memory == BRILLE_MEM_HOST) {
    if (q_out) HIP_TRY(hipMemcpy(q_out, dq, n * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (tau_out) HIP_TRY(hipMemcpy(tau_out, dtau, n * 3 * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (rot_out) HIP_TRY(hipMemcpy(rot_out, drot, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (invrot_out) HIP_TRY(hipMemcpy(invrot_out, dinv, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (status_out) HIP_TRY(hipMemcpy(status_out, dst, n, hipMemcpyDeviceToHost));
    free_all(tmp);
  }
  if (e.count) return report_point_error(e, Q, n, memory == BRILLE_MEM_HOST);
  return BRILLE_OK;
}

// ---- device plumbing -------------------------------------------------------
int brille_device_alloc(uint64_t bytes, void** ptr) {
  if (!ptr) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
  return BRILLE_OK;
}
int brille_device_free(void* ptr) {
  HIP_TRY(hipFree(ptr));
  return BRILLE_OK;
}
int brille_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return BRILLE_OK;
}
int brille_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return BRILLE_OK;
}
int brille_stream_create(void** stream) {
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return BRILLE_OK;
}
int brille_stream_destroy(void* stream) {
  HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return BRILLE_OK;
}
int brille_stream_sync(void* stream) {
  HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return BRILLE_OK;
}
int brille_device_sync(void) {
  HIP_TRY(hipDeviceSynchronize());
  return BRILLE_OK;
}

}  // extern "C"
