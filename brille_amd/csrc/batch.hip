// Many grids in one launch (no reference counterpart).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "grid_state.hpp"

using namespace brille_amd;

struct brille_batch {
  std::vector<brille_grid*> grids;
};

extern "C" {

int brille_batch_create(brille_grid* const* grids, uint32_t n_grids, brille_batch** out) {
  if (!grids || !out || n_grids == 0) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  auto* b = new brille_batch();
  for (uint32_t i = 0; i < n_grids; ++i) {
    if (!grids[i] || !grids[i]->filled) {
      delete b;
      return fail(BRILLE_ERR_NOT_FILLED, "The interpolation data must be filled before interpolating.");
    }
    b->grids.push_back(grids[i]);
  }
  *out = b;
  return BRILLE_OK;
}

int brille_batch_run(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts* opts,
                     void* const* values_out, void* const* vectors_out, void* stream) {
  if (!b || !offsets || !values_out || !vectors_out) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null argument");
  brille_run_opts o{};
  o.check = 1;
  if (opts) o = *opts;
  if (o.memory != BRILLE_MEM_DEVICE) return fail(BRILLE_ERR_INVALID_ARGUMENT, "batches take device memory");
  for (size_t g = 0; g < b->grids.size(); ++g) {
    const uint64_t n = offsets[g + 1] - offsets[g];
    int rc = run_interpolate(b->grids[g], Q + 3 * offsets[g], n, &o, values_out[g], vectors_out[g], nullptr,
                             static_cast<hipStream_t>(stream), nullptr, 0);
    if (rc != BRILLE_OK) return rc;
  }
  return BRILLE_OK;
}

int brille_batch_autotune(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts* opts,
                          void* const* values_out, void* const* vectors_out, void* stream) {
  return brille_batch_run(b, Q, offsets, opts, values_out, vectors_out, stream);
}

int brille_batch_schedule(const brille_batch* b, uint32_t* points_per_lane, uint32_t* block_size, uint32_t* tile_points) {
  if (!b) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null batch");
  if (points_per_lane) *points_per_lane = 1;
  if (block_size) *block_size = TILE;
  if (tile_points) *tile_points = TILE;
  return BRILLE_OK;
}

int brille_batch_destroy(brille_batch* b) {
  delete b;
  return BRILLE_OK;
}

int brille_time_batch_run(brille_batch* b, const double* Q, const uint64_t* offsets, const brille_run_opts* opts,
                          void* const* values_out, void* const* vectors_out, uint32_t iters, void* stream,
                          brille_kernel_times* times) {
  if (!times) return fail(BRILLE_ERR_INVALID_ARGUMENT, "null times");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(BRILLE_ERR_DEVICE, "hipEventCreate failed");
  brille_run_opts o{};
  if (opts) o = *opts;
  o.check = 0;
  (void)hipEventRecord(e0, s);
  for (uint32_t it = 0; it < (iters ? iters : 1); ++it) {
    int rc = brille_batch_run(b, Q, offsets, &o, values_out, vectors_out, stream);
    if (rc != BRILLE_OK) return rc;
  }
  (void)hipEventRecord(e1, s);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  times->n_kernels = 1;
  std::snprintf(times->names[0], sizeof(times->names[0]), "k_ir_interpolate");
  times->ms[0] = ms / (float)(iters ? iters : 1);
  times->total_ms = times->ms[0];
  return BRILLE_OK;
}

}  // extern "C"
