// Launch logic behind brille_ir_interpolate_at: validation, staging of host
// buffers, kernel selection, error reduction and event timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fast_tables.hpp"
#include "grid_state.hpp"

using namespace brille_amd;

namespace brille_amd {

#define HIP_TRY(expr)                                                                                            \
  do {                                                                                                           \
    hipError_t e_ = (expr);                                                                                      \
    if (e_ != hipSuccess) return fail(BRILLE_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));       \
  } while (0)

__global__ void k_err_init2(ErrRec* e) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    e->first_bad = ~0ull;
    e->count = 0ull;
  }
}

void free_staging(brille_grid* G) {
  if (G->stage) (void)hipFree(G->stage);
  G->stage = nullptr;
  G->stage_bytes = 0;
}

// Interpolator::rotate_in_place dispatch  src/interpolator.hpp:387-428
int check_rotation_modes(const brille_grid* G) {
  const DataConst& d = G->d;
  auto one = [&](int rot, int lu, bool cplx, bool has_vm, const char* what) -> int {
    switch (lu) {
      case BRILLE_LU_REAL_LATTICE:
        if (rot == BRILLE_ROT_VECTOR || rot == BRILLE_ROT_PSEUDOVECTOR) return BRILLE_OK;
        if (rot == BRILLE_ROT_GAMMA) {
          if (!cplx) return fail(BRILLE_ERR_ROTATION, "RotatesLike == Gamma requires complex valued data!");
          return BRILLE_OK;
        }
        return fail(BRILLE_ERR_ROTATION, "Impossible RotatesLike value!");
      case BRILLE_LU_RECIPROCAL_LATTICE:
        if (rot == BRILLE_ROT_VECTOR) return BRILLE_OK;
        break;
      case BRILLE_LU_ANGSTROM:
        if (rot == BRILLE_ROT_GAMMA) {
          if (!cplx) return fail(BRILLE_ERR_ROTATION, "RotatesLike == Gamma requires complex valued data!");
          return BRILLE_OK;
        }
        break;
      default:
        break;
    }
    (void)has_vm;
    return fail(BRILLE_ERR_ROTATION, std::string("LengthUnit, RotatesLike combination not implemented (") + what + ")");
  };
  int rc = one(d.val_rot, d.val_lu, false, d.ev[1] + d.ev[2] > 0, "values");
  if (rc != BRILLE_OK) return rc;
  rc = one(d.vec_rot, d.vec_lu, true, d.ec[1] + d.ec[2] > 0, "vectors");
  if (rc != BRILLE_OK) return rc;
  if (d.vec_rot == BRILLE_ROT_GAMMA && d.ec[1] + d.ec[2] > 0) {
    if (!d.F0) return fail(BRILLE_ERR_ROTATION, "Gamma rotation requires a GammaTable");
    if (d.ec[1] / 3u != d.n_atoms)
      return fail(BRILLE_ERR_ROTATION, "the eigenvectors must hold one 3-vector per GammaTable atom");
    if (d.ec[2] > 0) return fail(BRILLE_ERR_UNSUPPORTED, "Gamma transformation of 3N x 3N tensors is not supported");
    if (d.vec_lu == BRILLE_LU_ANGSTROM && !d.cart_rot) return fail(BRILLE_ERR_ROTATION, "Cartesian Gamma rotation needs Cartesian matrices");
  }
  return BRILLE_OK;
}

// brille_grid_fill_ex may keep one precision of the payload only
int check_payload_precision(const brille_grid* G, int precision) {
  if (precision == BRILLE_FP64 && !G->d.rows64)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "the payload was filled without its FP64 rows");
  if (precision == BRILLE_FP32 && !G->d.rows32)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "the payload was filled without its FP32 rows");
  return BRILLE_OK;
}

using TileKernel = void (*)(GridConst, DataConst, LaunchArgs);

// which specialised phase 2 the payload admits
int payload_layout(const DataConst& d) {
  const bool scalar_values = d.ev[1] == 0 && d.ev[2] == 0;
  if (d.n_keys || !scalar_values) return LAYOUT_GENERIC;
  if (d.ec[1] == 0 && d.ec[2] == 0) return LAYOUT_SCALAR;
  if (d.ec[0] == 0 && d.ec[2] == 0 && d.vec_rot == BRILLE_ROT_GAMMA &&
      (d.vec_lu == BRILLE_LU_REAL_LATTICE || d.vec_lu == BRILLE_LU_ANGSTROM))
    return LAYOUT_GAMMA;
  return LAYOUT_GENERIC;
}

// Gamma or scalar payloads whose rows are too long for one lane (more than 12
// and at most kWideMax complex elements, at most 64 values): one wave per row
bool wide_rows(const DataConst& d) {
  const int L = payload_layout(d);
  return (L == LAYOUT_GAMMA || L == LAYOUT_SCALAR) && d.sc > 12 && d.sc <= kWideMax && d.sv <= 64;
}

// The instantiated specialisations: Gamma payloads of 1-4 atoms and scalar
// payloads of 1-4 complex elements, each with one value per mode; anything
// else runs the general phase 2.
template <typename Real>
static TileKernel pick_tuned(const DataConst& d) {
  const int L = payload_layout(d);
  if (L == LAYOUT_GAMMA && d.sv == 1) {
    switch (d.sc) {
      case 3: return k_ir_interpolate<Real, true, LAYOUT_GAMMA, 3, 1>;
      case 6: return k_ir_interpolate<Real, true, LAYOUT_GAMMA, 6, 1>;
      case 9: return k_ir_interpolate<Real, true, LAYOUT_GAMMA, 9, 1>;
      case 12: return k_ir_interpolate<Real, true, LAYOUT_GAMMA, 12, 1>;
      default: break;
    }
  }
  if (L == LAYOUT_SCALAR && d.sv == 1) {
    switch (d.sc) {
      case 1: return k_ir_interpolate<Real, true, LAYOUT_SCALAR, 1, 1>;
      case 2: return k_ir_interpolate<Real, true, LAYOUT_SCALAR, 2, 1>;
      case 3: return k_ir_interpolate<Real, true, LAYOUT_SCALAR, 3, 1>;
      case 4: return k_ir_interpolate<Real, true, LAYOUT_SCALAR, 4, 1>;
      default: break;
    }
  }
  if (wide_rows(d)) return k_ir_interpolate<Real, true, LAYOUT_WIDE, 0, 0>;
  return k_ir_interpolate<Real, true, LAYOUT_GENERIC, 0, 0>;
}

// plain: the literal FP64 restatement; tuned: fast fold-in and specialised sums
static void launch_sel(const brille_grid* G, const LaunchArgs& a, int precision, bool fast, hipStream_t s) {
  const uint64_t blocks = (a.n + TILE - 1) / TILE;
  if (!blocks) return;
  TileKernel k;
  if (precision == BRILLE_FP32) k = fast ? pick_tuned<float>(G->d) : k_ir_interpolate<float, false, LAYOUT_GENERIC, 0, 0>;
  else k = fast ? pick_tuned<double>(G->d) : k_ir_interpolate<double, false, LAYOUT_GENERIC, 0, 0>;
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(TILE), 0, s, G->g, G->d, a);
}

void launch_tile(const brille_grid* G, const LaunchArgs& a, int precision, bool fast, hipStream_t s) {
  launch_sel(G, a, precision, fast, s);
}

// every kernel of one call on device buffers
static int dispatch(brille_grid* G, const LaunchArgs& a, const brille_run_opts& o, hipStream_t s) {
  const bool fast = o.kernel != BRILLE_KERNEL_PLAIN;
  if (fast && binned_applies(G, a, o.kernel)) return launch_binned(G, a, o.precision, s, nullptr);
  launch_sel(G, a, o.precision, fast, s);
  return BRILLE_OK;
}

int report_point_error(const ErrRec& e, const double* Q, uint64_t n, bool host_mem) {
  const uint64_t idx = e.first_bad & ((1ull << kErrStatusShift) - 1ull);
  const int st = (int)(e.first_bad >> kErrStatusShift);
  double row[3];
  const double* rp = nullptr;
  if (idx < n) {
    if (host_mem) rp = Q + 3 * idx;
    else if (hipMemcpy(row, Q + 3 * idx, sizeof(row), hipMemcpyDeviceToHost) == hipSuccess) rp = row;
  }
  char buf[200];
  std::string q = "#" + std::to_string(idx);
  if (rp) {
    std::snprintf(buf, sizeof(buf), "[%.17g, %.17g, %.17g]", rp[0], rp[1], rp[2]);
    q = buf;
  }
  switch (st) {
    case BRILLE_PT_OUTSIDE_BZ:
      return fail(BRILLE_ERR_OUTSIDE_BZ, "Not all points inside Brillouin zone (first: Q = " + q + ")");
    case BRILLE_PT_OUTSIDE_IR:
      return fail(BRILLE_ERR_OUTSIDE_IR, "Q = " + q + " is outside of the irreducible BrillouinZone");
    case BRILLE_PT_NULL_NODE:
      return fail(BRILLE_ERR_NULL_NODE, "Attempting to access null node for point " + q);
    default: {
      std::string m = "interpolate_at failed to find " + std::to_string(e.count) + " point";
      m += e.count > 1 ? "s." : ".";
      return fail(BRILLE_ERR_NOT_FOUND, m);
    }
  }
}

// proposals tried / verified by the call that just ran on stream s
static int read_proposal_stats(brille_grid* G, hipStream_t s) {
  unsigned long long st[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(st, G->prop_stats, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  G->last_tried = st[0];
  G->last_hits = st[1];
  return BRILLE_OK;
}

int run_interpolate(brille_grid* G, const double* Q, uint64_t n, const brille_run_opts* opts_in, void* values_out,
                    void* vectors_out, const brille_aux_out* aux_in, hipStream_t s, brille_kernel_times* times,
                    uint32_t iters, int fold) {
  if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
  if (!G->filled) return fail(BRILLE_ERR_NOT_FILLED, "The interpolation data must be filled before interpolating.");
  if (G->g.n_nodes == 0) return fail(BRILLE_ERR_INVALID_ARGUMENT, "this handle holds no trellis");
  brille_run_opts opts{};
  opts.check = 1;
  if (opts_in) opts = *opts_in;
  if (opts.precision != BRILLE_FP64 && opts.precision != BRILLE_FP32)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "unknown precision");
  int rc = check_rotation_modes(G);
  if (rc != BRILLE_OK) return rc;
  if (rc = check_payload_precision(G, opts.precision); rc != BRILLE_OK) return rc;
  if (n == 0) return BRILLE_OK;
  if (!Q || !values_out || !vectors_out) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null input or output pointer");
  if (opts.use_proposal && !G->prop_trained)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "use_proposal requires a trained proposal (brille_proposal_train)");
  if (opts.use_proposal && (opts.kernel == BRILLE_KERNEL_PLAIN || opts.kernel == BRILLE_KERNEL_TILE))
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "the learned proposal runs in the binned pipeline (kernel AUTO, TUNED or BINNED)");
  if (opts.use_proposal && aux_in && aux_in->tau)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "the binned pipeline does not report tau");
  const bool fast = opts.kernel != BRILLE_KERNEL_PLAIN;
  const size_t rs = opts.precision == BRILLE_FP32 ? 4 : 8;
  const size_t vbytes = (size_t)G->d.B * G->d.sv * rs;
  const size_t cbytes = (size_t)G->d.B * G->d.sc * 2 * rs;
  brille_aux_out aux{};
  if (aux_in) aux = *aux_in;
  LaunchArgs a{};
  a.fold = opts.no_move ? FOLD_NONE : fold;
  a.use_proposal = opts.use_proposal;
  a.prop = G->prop;
  if (opts.use_proposal) zero_words(G->prop_stats, 4, s);  // tried, verified (two 64-bit counters)
  a.err = G->err;
  if (opts.memory == BRILLE_MEM_DEVICE) {
    a.Q = Q;
    a.n = n;
    a.base = 0;
    a.vals = values_out;
    a.vecs = vectors_out;
    a.aux = aux;
    hipLaunchKernelGGL(k_err_init2, dim3(1), dim3(1), 0, s, G->err);
    if (times) {
      std::memset(times, 0, sizeof(*times));
      const uint32_t it_n = std::max(iters, 1u);
      if (fast && binned_applies(G, a, opts.kernel)) {
        // stages from events on the streams that ran them; the call from
        // events on `s` around all of it
        float st[3] = {0.f, 0.f, 0.f};
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, s));
        for (uint32_t it = 0; it < it_n && rc == BRILLE_OK; ++it) rc = launch_binned(G, a, opts.precision, s, st);
        if (rc == BRILLE_OK) {
          HIP_TRY(hipEventRecord(e1, s));
          HIP_TRY(hipEventSynchronize(e1));
          float ms = 0;
          HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
          times->total_ms = ms / (float)it_n;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != BRILLE_OK) return rc;
        const char* nm[3] = {"k_bin_fold", opts.use_proposal ? "k_bin_bucket+propose" : "k_bin_bucket", "k_bin_sums"};
        times->n_kernels = 3;
        for (int k = 0; k < 3; ++k) {
          std::snprintf(times->names[k], sizeof(times->names[k]), "%s", nm[k]);
          times->ms[k] = st[k] / (float)it_n;
        }
      } else {
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, s));
        for (uint32_t it = 0; it < it_n; ++it) launch_sel(G, a, opts.precision, fast, s);
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        times->n_kernels = 1;
        std::snprintf(times->names[0], sizeof(times->names[0]), "k_ir_interpolate");
        times->ms[0] = ms / (float)it_n;
        times->total_ms = times->ms[0];
      }
    } else {
      rc = dispatch(G, a, opts, s);
      if (rc != BRILLE_OK) return rc;
    }
    HIP_TRY(hipGetLastError());
    if (opts.use_proposal) {
      rc = read_proposal_stats(G, s);
      if (rc != BRILLE_OK) return rc;
    }
    if (!opts.check && !times) return BRILLE_OK;
    ErrRec e;
    HIP_TRY(hipMemcpyAsync(&e, G->err, sizeof(e), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (e.count) return report_point_error(e, Q, n, false);
    return BRILLE_OK;
  }
  // host memory: stream through device staging in chunks
  const uint64_t chunk = std::min<uint64_t>(n, 1ull << 20);
  struct Slot { void* host; size_t per; size_t off; };
  std::vector<Slot> slots;
  size_t off = 0;
  auto add = [&](void* h, size_t per) {
    slots.push_back({h, per, off});
    off += (per * chunk + 255) & ~size_t(255);
  };
  add(const_cast<double*>(Q), 3 * sizeof(double));
  add(values_out, vbytes);
  add(vectors_out, cbytes);
  void** auxp[10] = {(void**)&aux.ir_q, (void**)&aux.tau, (void**)&aux.rot, (void**)&aux.invrot, (void**)&aux.node,
                     (void**)&aux.tet, (void**)&aux.n_iw, (void**)&aux.iw_index, (void**)&aux.iw_weight,
                     (void**)&aux.status};
  const size_t auxper[10] = {24, 12, 4, 4, 4, 4, 1, 32, 64, 1};
  for (int i = 0; i < 10; ++i)
    if (*auxp[i]) add(*auxp[i], auxper[i]);
  if (G->stage_bytes < off) {
    free_staging(G);
    HIP_TRY(hipMalloc(&G->stage, off));
    G->stage_bytes = off;
  }
  char* st = static_cast<char*>(G->stage);
  hipLaunchKernelGGL(k_err_init2, dim3(1), dim3(1), 0, s, G->err);
  for (uint64_t c0 = 0; c0 < n; c0 += chunk) {
    const uint64_t m = std::min(chunk, n - c0);
    HIP_TRY(hipMemcpyAsync(st + slots[0].off, Q + 3 * c0, m * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    a.Q = reinterpret_cast<const double*>(st + slots[0].off);
    a.n = m;
    a.base = c0;
    a.vals = st + slots[1].off;
    a.vecs = st + slots[2].off;
    brille_aux_out da{};
    size_t k = 3;
    void** dauxp[10] = {(void**)&da.ir_q, (void**)&da.tau, (void**)&da.rot, (void**)&da.invrot, (void**)&da.node,
                        (void**)&da.tet, (void**)&da.n_iw, (void**)&da.iw_index, (void**)&da.iw_weight,
                        (void**)&da.status};
    for (int i = 0; i < 10; ++i)
      if (*auxp[i]) *dauxp[i] = st + slots[k++].off;
    a.aux = da;
    rc = dispatch(G, a, opts, s);
    if (rc != BRILLE_OK) return rc;
    HIP_TRY(hipGetLastError());
    for (size_t i = 1; i < slots.size(); ++i)
      HIP_TRY(hipMemcpyAsync(static_cast<char*>(slots[i].host) + c0 * slots[i].per, st + slots[i].off, m * slots[i].per,
                             hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  if (opts.use_proposal) {
    rc = read_proposal_stats(G, s);
    if (rc != BRILLE_OK) return rc;
  }
  ErrRec e;
  HIP_TRY(hipMemcpyAsync(&e, G->err, sizeof(e), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (e.count) return report_point_error(e, Q, n, true);
  return BRILLE_OK;
}

// Affine forms of the zone-face orientation tests and the seed orbit of the
// wedge (see fast_fold.hpp).  pseudo_orient3d(a,b,c,x) in a lattice with
// reciprocal volume V equals V det[a-x, b-x, c-x] on the components, i.e.
// V det[a,b,c] - x . V (a x b + b x c + c x a).
int build_fast_tables(const brille_grid_desc* d, brille_grid* G) {
  GridConst& g = G->g;
  FastTables ft = compute_fast_tables(d);
  g.aff_margin = ft.aff_margin;
  auto up = [&](const std::vector<double>& v, const double** dst) -> int {
    void* p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(double)) != hipSuccess) return fail(BRILLE_ERR_DEVICE, "hipMalloc failed");
    G->grid_allocs.push_back(p);
    if (hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(BRILLE_ERR_DEVICE, "hipMemcpy failed");
    *dst = static_cast<const double*>(p);
    return BRILLE_OK;
  };
  int rc = up(ft.prim_aff, &g.prim_aff);
  if (rc == BRILLE_OK) rc = up(ft.outer_aff, &g.outer_aff);
  if (rc != BRILLE_OK) return rc;
  if (ft.has_orbit) {
    rc = up(ft.orbit, &g.orbit);
    if (rc != BRILLE_OK) return rc;
    g.orbit_margin = ft.orbit_margin;
  } else {
    g.has_seed = 0;
  }
  return BRILLE_OK;
}


}  // namespace brille_amd

extern "C" {

int brille_time_ir_interpolate_at(brille_grid* G, const double* Q, uint64_t n, const brille_run_opts* opts,
                                  void* values_out, void* vectors_out, uint32_t iters, void* stream,
                                  brille_kernel_times* times) {
  BRILLE_LOCK_GRID(G);
  if (!times) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null times");
  if (opts && opts->memory != BRILLE_MEM_DEVICE) return fail(BRILLE_ERR_INVALID_ARGUMENT, "timing needs device memory");
  return run_interpolate(G, Q, n, opts, values_out, vectors_out, nullptr, static_cast<hipStream_t>(stream), times, iters);
}


}  // extern "C"
