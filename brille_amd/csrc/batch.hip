// Many grids in one persistent, graph-captured launch (no reference
// counterpart; the reference interpolates one grid per call).
//
// Every grid keeps its tables where brille_grid_create put them; the batch
// adds a device array of BatchGrid (the grid's GridConst/DataConst, its error
// record and which phase 2 its payload takes) and, per run, a device array of
// BatchSlot (where the grid's points and outputs are, and its first work
// item).  Work items are tiles of TB consecutive points of one grid, numbered
// slot after slot.  k_batch is persistent: each block takes `grab` tiles at a
// time from a device counter until none are left, finds the slot of each tile
// by binary search, and runs the tile's two phases with that grid's tables.
// k_batch_reset (a kernel, so that a captured graph replays it correctly)
// clears the counter and the per-grid error records.  Both launches are
// captured once per (precision, schedule) into a HIP graph; a run uploads the
// slots and replays the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "grid_state.hpp"

using namespace brille_amd;

namespace brille_amd {

enum BatchKind : int32_t {
  BK_GENERIC = 0,
  BK_GAMMA3 = 1,
  BK_GAMMA6 = 2,
  BK_SCALAR1 = 3,
  BK_SCALAR2 = 4,
  BK_SCALAR3 = 5,
  BK_SCALAR4 = 6,
  BK_WIDE = 7
};

struct BatchGrid {
  GridConst g;
  DataConst d;
  ErrRec* err;
  int32_t kind;
  int32_t ph_in_lds;
  uint32_t n_atoms;
  uint32_t pad;
};

struct BatchSlot {
  const double* Q;
  void* vals;
  void* vecs;
  uint64_t n;      // points of the grid
  uint64_t tile0;  // first work item of the grid
  uint32_t grid;   // index into the BatchGrid array
  uint32_t pad;
};

struct BatchHeader {
  unsigned long long counter;
  uint64_t n_tiles;
  uint32_t n_slots;
  int32_t fold;  // Fold
};

// the specialised phase 2 a grid's payload takes inside the batch kernel
static int32_t batch_kind(const DataConst& d) {
  const int L = payload_layout(d);
  if (L == LAYOUT_GAMMA && d.sv == 1 && d.sc == 3) return BK_GAMMA3;
  if (L == LAYOUT_GAMMA && d.sv == 1 && d.sc == 6) return BK_GAMMA6;
  if (L == LAYOUT_SCALAR && d.sv == 1 && d.sc >= 1 && d.sc <= 4) return BK_SCALAR1 + (int32_t)d.sc - 1;
  if (wide_rows(d)) return BK_WIDE;
  return BK_GENERIC;
}

template <typename Real, int L, int SC, int SV, int TB>
__device__ __forceinline__ void batch_tile(const BatchGrid& bg, const BatchSlot& sl, uint64_t p0, uint32_t npts,
                                           PointRec<Real>* s_rec, typename Cplx<Real>::T* s_ph, double* s_q,
                                           typename Cplx<Real>::T* s_wide, int fold) {
  const uint32_t tid = threadIdx.x;
  PointRec<Real>& r = s_rec[tid];
  r.nslot = 0;
  if (tid < npts) {
    const uint64_t p = p0 + tid;
    double Qp[3] = {sl.Q[3 * p], sl.Q[3 * p + 1], sl.Q[3 * p + 2]};
    Located P;
    locate<true>(bg.g, Qp, fold, P);
    if (P.status != BRILLE_PT_OK) report_point(bg.err, p, P.status);
    else store_record<Real, L, SC>(bg.g, bg.d, P, r, s_ph, s_q, tid, bg.ph_in_lds != 0, bg.n_atoms);
  }
  __syncthreads();
  // four rows in flight: the persistent kernel's registers are the largest of all its branches
  tile_sums<Real, true, L, SC, SV, TB, 4>(bg.g, bg.d, s_rec, s_ph, s_q, nullptr, p0, npts, bg.ph_in_lds != 0, sl.vals,
                                         sl.vecs, s_wide);
  __syncthreads();
}

// FP32: three waves per SIMD (the wide-row branch needs the occupancy; one
// spilled dword)
template <typename Real, int TB>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 3 : 1)))
void k_batch(const BatchGrid* grids, const BatchSlot* slots, BatchHeader* hdr, uint32_t grab) {
  using C = typename Cplx<Real>::T;
  __shared__ PointRec<Real> s_rec[TB];
  __shared__ C s_ph[TB * 4];
  // q (general and wide layouts) or the Gamma frames
  __shared__ double s_q[(3 > frame_words<Real>() ? 3 : frame_words<Real>()) * TB];
  __shared__ C s_wide[(TB / 64) * kWideMax];
  __shared__ unsigned long long s_t;
  const uint64_t n_tiles = hdr->n_tiles;
  const uint32_t n_slots = hdr->n_slots;
  const int fold = hdr->fold;
  for (;;) {
    if (threadIdx.x == 0) s_t = atomicAdd(&hdr->counter, (unsigned long long)grab);
    __syncthreads();
    const uint64_t t0 = s_t;
    __syncthreads();
    if (t0 >= n_tiles) break;
    const uint64_t t1 = t0 + grab < n_tiles ? t0 + grab : n_tiles;
    for (uint64_t t = t0; t < t1; ++t) {
      uint32_t lo = 0, hi = n_slots;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (slots[mid].tile0 <= t) lo = mid;
        else hi = mid;
      }
      lo = __builtin_amdgcn_readfirstlane(lo);
      const BatchSlot& sl = slots[lo];
      const BatchGrid& bg = grids[sl.grid];
      const uint64_t p0 = (t - sl.tile0) * TB;
      const uint64_t left = sl.n - p0;
      const uint32_t npts = left < (uint64_t)TB ? (uint32_t)left : (uint32_t)TB;
      switch (bg.kind) {
        case BK_GAMMA3: batch_tile<Real, LAYOUT_GAMMA, 3, 1, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
        case BK_GAMMA6: batch_tile<Real, LAYOUT_GAMMA, 6, 1, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
        case BK_SCALAR1: batch_tile<Real, LAYOUT_SCALAR, 1, 1, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
        case BK_SCALAR2: batch_tile<Real, LAYOUT_SCALAR, 2, 1, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
        case BK_SCALAR3: batch_tile<Real, LAYOUT_SCALAR, 3, 1, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
        case BK_SCALAR4: batch_tile<Real, LAYOUT_SCALAR, 4, 1, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
        case BK_WIDE: batch_tile<Real, LAYOUT_WIDE, 0, 0, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
        default: batch_tile<Real, LAYOUT_GENERIC, 0, 0, TB>(bg, sl, p0, npts, s_rec, s_ph, s_q, s_wide, fold); break;
      }
    }
  }
}

// graph-safe clear of the work counter and of every grid's error record
__global__ void k_batch_reset(BatchHeader* hdr, const BatchGrid* grids, uint32_t n_grids) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) hdr->counter = 0ull;
  if (i < n_grids) {
    grids[i].err->first_bad = ~0ull;
    grids[i].err->count = 0ull;
  }
}

// a member grid as the batch kernel sees it
static void describe_batch_grid(const brille_grid* G, BatchGrid& bg) {
  std::memset(&bg, 0, sizeof(bg));
  bg.g = G->g;
  bg.d = G->d;
  bg.err = G->err;
  bg.kind = batch_kind(G->d);
  const bool want_phase = G->d.vec_rot == BRILLE_ROT_GAMMA && G->d.ec[1] > 0;
  bg.n_atoms = G->d.ec[1] / 3u;
  bg.ph_in_lds = (bg.kind == BK_GAMMA3 || bg.kind == BK_GAMMA6 || (want_phase && bg.n_atoms <= 4)) ? 1 : 0;
}

struct Schedule {
  uint32_t tile = 256;       // points per tile = threads per block
  uint32_t blocks_per_cu = 4;
  uint32_t grab = 1;
  uint32_t order = 1;        // 0: as given, 1: costliest points first
  // grids given at least this many points run their own binned pipeline
  // after the persistent kernel (their rows are then read in node order)
  uint64_t binned_min = ~0ull;
};

}  // namespace brille_amd

struct brille_batch {
  std::vector<brille_grid*> grids;
  BatchGrid* d_grids = nullptr;
  BatchSlot* d_slots = nullptr;
  BatchHeader* d_hdr = nullptr;
  uint32_t cap_slots = 0;
  uint32_t n_cu = 0;
  Schedule sched;
  hipStream_t cap_stream = nullptr;
  // streams the binned grids spread over beside the caller's (created on
  // first use, default priority), and the events that fork them from the
  // caller's stream and join them back
  static constexpr int kLanes = 3;
  hipStream_t lanes[kLanes] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kLanes] = {};
  // instantiated graphs by (precision, tile, blocks, grab)
  std::map<std::tuple<int, uint32_t, uint32_t, uint32_t>, hipGraphExec_t> graphs;
  // host copy of d_grids, refreshed when a grid's tables change, and the
  // fill generation of each grid that copy was made from
  std::vector<brille_amd::BatchGrid> host_grids;
  std::vector<uint64_t> fill_gen;
  // the member grids' locks, taken for every run (address order)
  std::vector<const brille_grid*> lock_set() const { return {grids.begin(), grids.end()}; }
};

namespace brille_amd {

#define BATCH_TRY(expr)                                                                             \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return fail(BRILLE_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename Real>
static void launch_batch_kernel(brille_batch* b, uint32_t tile, uint32_t blocks, uint32_t grab, hipStream_t s) {
  if (tile == 128)
    hipLaunchKernelGGL((k_batch<Real, 128>), dim3(blocks), dim3(128), 0, s, b->d_grids, b->d_slots, b->d_hdr, grab);
  else
    hipLaunchKernelGGL((k_batch<Real, 256>), dim3(blocks), dim3(256), 0, s, b->d_grids, b->d_slots, b->d_hdr, grab);
}

static int graph_for(brille_batch* b, int precision, const Schedule& sc, hipGraphExec_t* out) {
  const uint32_t blocks = std::max(1u, b->n_cu * sc.blocks_per_cu);
  auto key = std::make_tuple(precision, sc.tile, blocks, sc.grab);
  auto it = b->graphs.find(key);
  if (it != b->graphs.end()) {
    *out = it->second;
    return BRILLE_OK;
  }
  hipGraph_t graph;
  BATCH_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
  const uint32_t ng = (uint32_t)b->grids.size();
  hipLaunchKernelGGL(k_batch_reset, dim3((ng + 63) / 64), dim3(64), 0, b->cap_stream, b->d_hdr, b->d_grids, ng);
  if (precision == BRILLE_FP32) launch_batch_kernel<float>(b, sc.tile, blocks, sc.grab, b->cap_stream);
  else launch_batch_kernel<double>(b, sc.tile, blocks, sc.grab, b->cap_stream);
  BATCH_TRY(hipStreamEndCapture(b->cap_stream, &graph));
  hipGraphExec_t exec;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(BRILLE_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  b->graphs[key] = exec;
  *out = exec;
  return BRILLE_OK;
}

// per-point cost of a grid for the costliest-first order: payload reals per point
static double point_cost(const brille_grid* G) {
  const DataConst& d = G->d;
  return 64.0 + (double)d.B * (2.0 * d.sc + d.sv);
}

// slots of one run (empty grids take no slot) and their upload
static int stage_slots(brille_batch* b, const double* Q, const uint64_t* offsets, void* const* vals,
                       void* const* vecs, const Schedule& sc, int fold, hipStream_t s) {
  const uint32_t ng = (uint32_t)b->grids.size();
  std::vector<uint32_t> order(ng);
  for (uint32_t g = 0; g < ng; ++g) order[g] = g;
  if (sc.order == 1)
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return point_cost(b->grids[x]) > point_cost(b->grids[y]); });
  std::vector<BatchSlot> slots;
  uint64_t tiles = 0;
  for (uint32_t g : order) {
    const uint64_t n = offsets[g + 1] - offsets[g];
    if (n == 0 || n >= sc.binned_min) continue;
    BatchSlot sl{};
    sl.Q = Q + 3 * offsets[g];
    sl.vals = vals[g];
    sl.vecs = vecs[g];
    sl.n = n;
    sl.tile0 = tiles;
    sl.grid = g;
    slots.push_back(sl);
    tiles += (n + sc.tile - 1) / sc.tile;
  }
  BatchHeader h{};
  h.counter = 0;
  h.n_tiles = tiles;
  h.n_slots = (uint32_t)slots.size();
  h.fold = fold;
  if (!slots.empty())
    BATCH_TRY(hipMemcpyAsync(b->d_slots, slots.data(), slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, s));
  BATCH_TRY(hipMemcpyAsync(b->d_hdr, &h, sizeof(h), hipMemcpyHostToDevice, s));
  // the copies read host memory that goes out of scope: finish them first
  BATCH_TRY(hipStreamSynchronize(s));
  return BRILLE_OK;
}

static int check_args(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts& o,
                      void* const* vals, void* const* vecs) {
  if (!b || !offsets || !vals || !vecs) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  if (o.memory != BRILLE_MEM_DEVICE) return fail(BRILLE_ERR_INVALID_ARGUMENT, "batches take device memory");
  if (o.precision != BRILLE_FP64 && o.precision != BRILLE_FP32)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "unknown precision");
  if (o.use_proposal) return fail(BRILLE_ERR_UNSUPPORTED, "batches do not use the learned proposal");
  const uint32_t ng = (uint32_t)b->grids.size();
  for (uint32_t g = 0; g < ng; ++g) {
    if (offsets[g + 1] < offsets[g]) return fail(BRILLE_ERR_INVALID_ARGUMENT, "batch offsets must not decrease");
    if (offsets[g + 1] > offsets[g] && (!Q || !vals[g] || !vecs[g]))
      return fail(BRILLE_ERR_INVALID_ARGUMENT, "null input or output pointer");
    if (!b->grids[g]->filled)
      return fail(BRILLE_ERR_NOT_FILLED, "grid " + std::to_string(g) + ": The interpolation data must be filled before interpolating.");
    int rc = check_rotation_modes(b->grids[g]);
    if (rc == BRILLE_OK) rc = check_payload_precision(b->grids[g], o.precision);
    if (rc != BRILLE_OK) return fail(rc, "grid " + std::to_string(g) + ": " + std::string(g_last_error_text()));
  }
  return BRILLE_OK;
}

// the first failing point of the lowest-numbered grid that has one
static int batch_errors(brille_batch* b, const double* Q, const uint64_t* offsets, hipStream_t s) {
  BATCH_TRY(hipStreamSynchronize(s));
  for (size_t g = 0; g < b->grids.size(); ++g) {
    ErrRec e;
    BATCH_TRY(hipMemcpy(&e, b->grids[g]->err, sizeof(e), hipMemcpyDeviceToHost));
    if (e.count) {
      const uint64_t n = offsets[g + 1] - offsets[g];
      int rc = report_point_error(e, Q + 3 * offsets[g], n, false);
      return fail(rc, "grid " + std::to_string(g) + ": " + std::string(g_last_error_text()));
    }
  }
  return BRILLE_OK;
}

static int run_sequential(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts& o,
                          void* const* vals, void* const* vecs, hipStream_t s) {
  for (size_t g = 0; g < b->grids.size(); ++g) {
    const uint64_t n = offsets[g + 1] - offsets[g];
    int rc = run_interpolate(b->grids[g], Q + 3 * offsets[g], n, &o, vals[g], vecs[g], nullptr, s, nullptr, 0);
    if (rc != BRILLE_OK) return fail(rc, "grid " + std::to_string(g) + ": " + std::string(g_last_error_text()));
  }
  return BRILLE_OK;
}

// The wide factor tables of the wide-row grids (built on first use, the same
// tables their own binned calls use); the grid copies on the device follow.
// A member refilled since the batch last copied its descriptors (its payload
// buffers were replaced) is described again; an unfilled one stops the run.
static int ensure_batch_tables(brille_batch* b, int precision, hipStream_t s) {
  bool changed = false;
  for (size_t g = 0; g < b->grids.size(); ++g) {
    brille_grid* G = b->grids[g];
    if (!G->filled) return fail(BRILLE_ERR_NOT_FILLED, "grid " + std::to_string(g) + ": The interpolation data must be filled before interpolating.");
    if (b->fill_gen[g] != G->fill_gen) {
      describe_batch_grid(G, b->host_grids[g]);
      b->fill_gen[g] = G->fill_gen;
      changed = true;
    }
    if (!wide_rows(G->d)) continue;
    const void* before = precision == BRILLE_FP32 ? (const void*)G->d.wide_fac32 : (const void*)G->d.wide_fac64;
    const int rc = ensure_cube_rows(G, precision, s);
    if (rc != BRILLE_OK) return rc;
    const void* after = precision == BRILLE_FP32 ? (const void*)G->d.wide_fac32 : (const void*)G->d.wide_fac64;
    if (after != before) {
      b->host_grids[g].d = G->d;
      changed = true;
    }
  }
  if (changed) {
    BATCH_TRY(hipStreamSynchronize(s));
    BATCH_TRY(hipMemcpy(b->d_grids, b->host_grids.data(), b->host_grids.size() * sizeof(BatchGrid),
                        hipMemcpyHostToDevice));
  }
  return BRILLE_OK;
}

static int launch_binned_grids(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts& o,
                               void* const* vals, void* const* vecs, hipStream_t s, const Schedule& sc);

static int run_with(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts& o,
                    void* const* vals, void* const* vecs, hipStream_t s, const Schedule& sc) {
  int rc = ensure_batch_tables(b, o.precision, s);
  if (rc != BRILLE_OK) return rc;
  rc = stage_slots(b, Q, offsets, vals, vecs, sc, o.no_move ? FOLD_NONE : FOLD_IR, s);
  if (rc != BRILLE_OK) return rc;
  hipGraphExec_t exec;
  rc = graph_for(b, o.precision, sc, &exec);
  if (rc != BRILLE_OK) return rc;
  BATCH_TRY(hipGraphLaunch(exec, s));
  return launch_binned_grids(b, Q, offsets, o, vals, vecs, s, sc);
}

// The large grids after the graph (whose reset clears every grid's error
// record), each a binned call; errors are read by the caller as for the
// others.  One grid's call leaves most of the GPU idle at its start and end
// (and all of it but a few blocks for a grid of few tiles), so the calls
// spread over the caller's stream and the batch's lanes, largest first onto
// the least loaded (by output bytes), and the caller's stream waits for all.
static int launch_binned_grids(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts& o,
                               void* const* vals, void* const* vecs, hipStream_t s, const Schedule& sc) {
  brille_run_opts ob = o;
  ob.kernel = BRILLE_KERNEL_BINNED;
  ob.check = 0;
  std::vector<std::pair<double, uint32_t>> big;  // (output bytes, grid)
  for (size_t g = 0; g < b->grids.size(); ++g) {
    const uint64_t n = offsets[g + 1] - offsets[g];
    if (n == 0 || n < sc.binned_min) continue;
    const DataConst& d = b->grids[g]->d;
    big.push_back({(double)n * d.B * (d.sv + 2.0 * d.sc), (uint32_t)g});
  }
  if (big.empty()) return BRILLE_OK;
  std::stable_sort(big.begin(), big.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  const int n_lanes = big.size() > 1 ? brille_batch::kLanes : 0;
  if (n_lanes) {
    for (int k = 0; k < n_lanes; ++k)
      if (!b->lanes[k]) {
        BATCH_TRY(hipStreamCreateWithFlags(&b->lanes[k], hipStreamNonBlocking));
        BATCH_TRY(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming));
      }
    if (!b->fork) BATCH_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    BATCH_TRY(hipEventRecord(b->fork, s));
    for (int k = 0; k < n_lanes; ++k) BATCH_TRY(hipStreamWaitEvent(b->lanes[k], b->fork, 0));
  }
  double load[brille_batch::kLanes + 1] = {};
  int rc = BRILLE_OK;
  for (const auto& [bytes, g] : big) {
    int k = 0;
    for (int j = 1; j <= n_lanes; ++j)
      if (load[j] < load[k]) k = j;
    load[k] += bytes;
    hipStream_t st = k == 0 ? s : b->lanes[k - 1];
    rc = run_interpolate(b->grids[g], Q + 3 * offsets[g], offsets[g + 1] - offsets[g], &ob, vals[g], vecs[g], nullptr,
                         st, nullptr, 0);
    if (rc != BRILLE_OK) {
      rc = fail(rc, "grid " + std::to_string(g) + ": " + std::string(g_last_error_text()));
      break;
    }
  }
  // the caller's stream waits for every lane it forked, also after an error
  for (int k = 0; k < n_lanes; ++k) {
    BATCH_TRY(hipEventRecord(b->join[k], b->lanes[k]));
    BATCH_TRY(hipStreamWaitEvent(s, b->join[k], 0));
  }
  return rc;
}

}  // namespace brille_amd

extern "C" {

int brille_batch_create(brille_grid* const* grids, uint32_t n_grids, brille_batch** out) {
  if (!grids || !out || n_grids == 0) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<BatchGrid> host(n_grids);
  for (uint32_t i = 0; i < n_grids; ++i) {
    brille_grid* G = grids[i];
    if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
    if (!G->filled) return fail(BRILLE_ERR_NOT_FILLED, "The interpolation data must be filled before interpolating.");
    if (G->g.n_nodes == 0) return fail(BRILLE_ERR_INVALID_ARGUMENT, "this handle holds no trellis");
    describe_batch_grid(G, host[i]);
  }
  auto* b = new brille_batch();
  b->grids.assign(grids, grids + n_grids);
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    delete b;
    return fail(BRILLE_ERR_DEVICE, "hipGetDeviceProperties failed");
  }
  b->n_cu = (uint32_t)prop.multiProcessorCount;
  void *pg = nullptr, *ps = nullptr, *ph = nullptr;
  if (hipMalloc(&pg, n_grids * sizeof(BatchGrid)) != hipSuccess || hipMalloc(&ps, n_grids * sizeof(BatchSlot)) != hipSuccess ||
      hipMalloc(&ph, sizeof(BatchHeader)) != hipSuccess ||
      hipMemcpy(pg, host.data(), n_grids * sizeof(BatchGrid), hipMemcpyHostToDevice) != hipSuccess ||
      hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) {
    if (pg) (void)hipFree(pg);
    if (ps) (void)hipFree(ps);
    if (ph) (void)hipFree(ph);
    delete b;
    return fail(BRILLE_ERR_DEVICE, "device allocation for the batch failed");
  }
  b->d_grids = static_cast<BatchGrid*>(pg);
  b->host_grids = host;
  b->fill_gen.resize(n_grids);
  for (uint32_t i = 0; i < n_grids; ++i) b->fill_gen[i] = grids[i]->fill_gen;
  b->d_slots = static_cast<BatchSlot*>(ps);
  b->d_hdr = static_cast<BatchHeader*>(ph);
  b->cap_slots = n_grids;
  *out = b;
  return BRILLE_OK;
}

int brille_batch_run(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts* opts,
                     void* const* values_out, void* const* vectors_out, void* stream) {
  if (!b) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null batch");
  GridLocks member_locks(b->lock_set());  // the member grids stay as they are for the whole run
  brille_run_opts o{};
  o.check = 1;
  if (opts) o = *opts;
  int rc = check_args(b, Q, offsets, o, values_out, vectors_out);
  if (rc != BRILLE_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (o.kernel == BRILLE_KERNEL_PLAIN || o.kernel == BRILLE_KERNEL_TILE)
    return run_sequential(b, Q, offsets, o, values_out, vectors_out, s);
  rc = run_with(b, Q, offsets, o, values_out, vectors_out, s, b->sched);
  if (rc != BRILLE_OK) return rc;
  if (!o.check) return BRILLE_OK;
  return batch_errors(b, Q, offsets, s);
}

int brille_batch_autotune(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts* opts,
                          void* const* values_out, void* const* vectors_out, void* stream) {
  if (!b) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null batch");
  GridLocks member_locks(b->lock_set());  // the member grids stay as they are for the whole run
  brille_run_opts o{};
  o.check = 1;
  if (opts) o = *opts;
  int rc = check_args(b, Q, offsets, o, values_out, vectors_out);
  if (rc != BRILLE_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  BATCH_TRY(hipEventCreate(&e0));
  BATCH_TRY(hipEventCreate(&e1));
  Schedule best = b->sched;
  float best_ms = -1.f;
  auto time_one = [&](const Schedule& sc, float* ms_out) -> int {
    int r = run_with(b, Q, offsets, o, values_out, vectors_out, s, sc);  // warm (and capture)
    if (r != BRILLE_OK) return r;
    float tot = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipEventRecord(e0, s);
      r = run_with(b, Q, offsets, o, values_out, vectors_out, s, sc);
      if (r != BRILLE_OK) return r;
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      tot += ms;
    }
    *ms_out = tot;
    return BRILLE_OK;
  };
  for (uint32_t tile : {128u, 256u})
    for (uint32_t bpc : {2u, 4u, 8u})
      for (uint32_t grab : {1u, 4u})
        for (uint32_t order : {0u, 1u}) {
          if (rc != BRILLE_OK) break;
          Schedule sc{tile, bpc, grab, order, ~0ull};
          float ms = 0.f;
          rc = time_one(sc, &ms);
          if (rc == BRILLE_OK && (best_ms < 0.f || ms < best_ms)) {
            best_ms = ms;
            best = sc;
          }
        }
  // then the share of the work the binned pipeline takes, with that schedule
  const Schedule persistent = best;
  for (uint64_t bmin : {1ull << 20, 1ull << 18, 1ull << 16, 1ull << 14, 1ull << 12, 1ull}) {
    if (rc != BRILLE_OK) break;
    Schedule sc = persistent;
    sc.binned_min = bmin;
    float ms = 0.f;
    rc = time_one(sc, &ms);
    if (rc == BRILLE_OK && ms < best_ms) {
      best_ms = ms;
      best = sc;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != BRILLE_OK) return rc;
  b->sched = best;
  rc = run_with(b, Q, offsets, o, values_out, vectors_out, s, best);
  if (rc != BRILLE_OK) return rc;
  return o.check ? batch_errors(b, Q, offsets, s) : BRILLE_OK;
}

int brille_batch_schedule(const brille_batch* b, uint32_t* tile_points, uint32_t* blocks, uint32_t* tiles_per_grab,
                          uint32_t* order, uint64_t* binned_min) {
  if (!b) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null batch");
  if (tile_points) *tile_points = b->sched.tile;
  if (blocks) *blocks = std::max(1u, b->n_cu * b->sched.blocks_per_cu);
  if (tiles_per_grab) *tiles_per_grab = b->sched.grab;
  if (order) *order = b->sched.order;
  if (binned_min) *binned_min = b->sched.binned_min;
  return BRILLE_OK;
}

int brille_batch_set_schedule(brille_batch* b, uint32_t tile_points, uint32_t blocks_per_cu, uint32_t tiles_per_grab,
                              uint32_t order, uint64_t binned_min) {
  if (!b) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null batch");
  if ((tile_points != 128 && tile_points != 256) || blocks_per_cu == 0 || blocks_per_cu > 64 || tiles_per_grab == 0 ||
      tiles_per_grab > 64 || order > 1)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "schedule out of range (tile 128 or 256, 1-64 blocks per CU, 1-64 tiles per grab, order 0 or 1)");
  GridLocks member_locks(b->lock_set());
  b->sched = Schedule{tile_points, blocks_per_cu, tiles_per_grab, order, binned_min};
  return BRILLE_OK;
}

// touches only what the batch owns: its member grids may already be gone
int brille_batch_destroy(brille_batch* b) {
  if (!b) return BRILLE_OK;
  (void)hipDeviceSynchronize();
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  if (b->d_grids) (void)hipFree(b->d_grids);
  if (b->d_slots) (void)hipFree(b->d_slots);
  if (b->d_hdr) (void)hipFree(b->d_hdr);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  for (int k = 0; k < brille_batch::kLanes; ++k) {
    if (b->lanes[k]) (void)hipStreamDestroy(b->lanes[k]);
    if (b->join[k]) (void)hipEventDestroy(b->join[k]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  delete b;
  return BRILLE_OK;
}

int brille_time_batch_run(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts* opts,
                          void* const* values_out, void* const* vectors_out, uint32_t iters, void* stream,
                          brille_kernel_times* times) {
  if (!b) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null batch");
  GridLocks member_locks(b->lock_set());  // the member grids stay as they are for the whole run
  if (!times) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null times");
  hipStream_t s = static_cast<hipStream_t>(stream);
  brille_run_opts o{};
  if (opts) o = *opts;
  o.check = 0;
  int rc = check_args(b, Q, offsets, o, values_out, vectors_out);
  if (rc != BRILLE_OK) return rc;
  const bool seq = o.kernel == BRILLE_KERNEL_PLAIN || o.kernel == BRILLE_KERNEL_TILE;
  // slots and graph outside the timed region: a replay is what is timed
  if (!seq) {
    rc = run_with(b, Q, offsets, o, values_out, vectors_out, s, b->sched);
    if (rc != BRILLE_OK) return rc;
  }
  hipGraphExec_t exec = nullptr;
  if (!seq) {
    rc = graph_for(b, o.precision, b->sched, &exec);
    if (rc != BRILLE_OK) return rc;
  }
  hipEvent_t e0, e1;
  BATCH_TRY(hipEventCreate(&e0));
  BATCH_TRY(hipEventCreate(&e1));
  BATCH_TRY(hipStreamSynchronize(s));
  (void)hipEventRecord(e0, s);
  for (uint32_t it = 0; it < (iters ? iters : 1); ++it) {
    if (seq) {
      rc = run_sequential(b, Q, offsets, o, values_out, vectors_out, s);
      if (rc != BRILLE_OK) return rc;
    } else {
      BATCH_TRY(hipGraphLaunch(exec, s));
      rc = launch_binned_grids(b, Q, offsets, o, values_out, vectors_out, s, b->sched);
      if (rc != BRILLE_OK) return rc;
    }
  }
  (void)hipEventRecord(e1, s);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::memset(times, 0, sizeof(*times));
  times->n_kernels = 1;
  std::snprintf(times->names[0], sizeof(times->names[0]), "%s", seq ? "per-grid calls" : "k_batch");
  times->ms[0] = ms / (float)(iters ? iters : 1);
  times->total_ms = times->ms[0];
  return BRILLE_OK;
}

}  // extern "C"
