// brille_ir_interpolate_at_noalloc: the hot path without any device
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// Everything brille_ir_interpolate_at keeps in device memory of its own for
// a call -- the error record, and for the binned pipeline the fold and sort
// records, bucket histograms, scans, node-order permutation and proposal list
// of a pass -- lives here in one file-scope device array, `scratch`, of
// kScratchBytes.  Each GPU holds its own copy at its own address from module
// load until the process ends; the first call on a device looks that address
// up (hipGetSymbolAddress) and keeps it.  The binned pipeline runs in passes
// sized to fit (2^19 points) on the caller's stream alone, with the row table
// only if an earlier call built it.  Inputs and outputs are caller device
// memory, as in brille_ir_interpolate_at with BRILLE_MEM_DEVICE.
//
// Every call on a GPU uses that GPU's array and leaves its contents there:
// the error record and the proposal counters are cleared by kernels at the
// start of each call, and every other buffer is written before it is read.
// Calls whose work overlaps on one GPU race on the array (issue them on one
// stream, or synchronise between them); within this process the host side of
// a call holds a per-device lock while it enqueues.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "grid_state.hpp"

namespace brille_amd {

constexpr uint64_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu[kMaxDevices];
static void* g_scratch_ptr[kMaxDevices];

// the device address of this device's `scratch` (looked up once per device)
static int scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(BRILLE_ERR_UNSUPPORTED, "device index beyond the scratch table");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BRILLE_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return BRILLE_OK;
}

__global__ void k_noalloc_err_init(ErrRec* e) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    e->first_bad = ~0ull;
    e->count = 0ull;
  }
}

}  // namespace brille_amd

using namespace brille_amd;

extern "C" int brille_ir_interpolate_at_noalloc(brille_grid* G, const double* Q, uint64_t n,
                                                const brille_run_opts* opts_in, void* values_out,
                                                void* vectors_out, const brille_aux_out* aux_in, void* stream) {
  BRILLE_LOCK_GRID(G);
  if (!G) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null grid");
  if (!G->filled) return fail(BRILLE_ERR_NOT_FILLED, "The interpolation data must be filled before interpolating.");
  if (G->g.n_nodes == 0) return fail(BRILLE_ERR_INVALID_ARGUMENT, "this handle holds no trellis");
  brille_run_opts o{};
  o.check = 1;
  if (opts_in) o = *opts_in;
  if (o.memory != BRILLE_MEM_DEVICE)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "the allocation-free call takes device memory only");
  if (o.precision != BRILLE_FP64 && o.precision != BRILLE_FP32) return fail(BRILLE_ERR_INVALID_ARGUMENT, "unknown precision");
  int rc = check_rotation_modes(G);
  if (rc == BRILLE_OK) rc = check_payload_precision(G, o.precision);
  if (rc != BRILLE_OK) return rc;
  if (n == 0) return BRILLE_OK;
  if (!Q || !values_out || !vectors_out) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null input or output pointer");
  if (o.use_proposal && !G->prop_trained)
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "use_proposal requires a trained proposal (brille_proposal_train)");
  if (o.use_proposal && (o.kernel == BRILLE_KERNEL_PLAIN || o.kernel == BRILLE_KERNEL_TILE))
    return fail(BRILLE_ERR_INVALID_ARGUMENT, "the learned proposal runs in the binned pipeline (kernel AUTO, TUNED or BINNED)");
  if (o.use_proposal && aux_in && aux_in->tau) return fail(BRILLE_ERR_INVALID_ARGUMENT, "the binned pipeline does not report tau");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(BRILLE_ERR_DEVICE, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return fail(BRILLE_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> device_lock(g_scratch_mu[dev]);
  void* base = nullptr;
  if ((rc = scratch_address(dev, &base)) != BRILLE_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  ErrRec* err = static_cast<ErrRec*>(base);
  void* ws = static_cast<char*>(base) + 256;
  const uint64_t ws_bytes = kScratchBytes - 256;
  LaunchArgs a{};
  a.Q = Q;
  a.n = n;
  a.base = 0;
  a.vals = values_out;
  a.vecs = vectors_out;
  if (aux_in) a.aux = *aux_in;
  a.err = err;
  a.fold = o.no_move ? FOLD_NONE : FOLD_IR;
  a.use_proposal = o.use_proposal;
  a.prop = G->prop;
  const bool fast = o.kernel != BRILLE_KERNEL_PLAIN;
  const bool binned = fast && binned_applies(G, a, o.kernel);
  uint64_t chunk = 0;
  if (binned) {
    chunk = binned_pass_for(G, ws_bytes);
    if (chunk == 0)
      return fail(BRILLE_ERR_UNSUPPORTED, "the binned pipeline of this grid needs more than the " +
                                              std::to_string(kScratchBytes >> 20) + " MiB of scratch per pass");
  }
  hipLaunchKernelGGL(k_noalloc_err_init, dim3(1), dim3(1), 0, s, err);
  if (o.use_proposal) zero_words(G->prop_stats, 4, s);
  if (binned) {
    rc = run_binned_passes(G, a, o.precision, s, nullptr, ws, chunk, 1);
    if (rc != BRILLE_OK) return rc;
  } else {
    launch_tile(G, a, o.precision, fast, s);
  }
  if (hipGetLastError() != hipSuccess) return fail(BRILLE_ERR_DEVICE, "launch failed");
  if (o.use_proposal) {
    unsigned long long st[2] = {0, 0};
    if (hipMemcpyAsync(st, G->prop_stats, sizeof(st), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(BRILLE_ERR_DEVICE, "reading the proposal counters failed");
    G->last_tried = st[0];
    G->last_hits = st[1];
  }
  if (!o.check) return BRILLE_OK;
  ErrRec e;
  if (hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "reading the error record failed");
  if (e.count) return report_point_error(e, Q, n, false);
  return BRILLE_OK;
}
