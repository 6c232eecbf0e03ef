// Host-side state behind the opaque brille_grid / brille_batch handles.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/brille_amd.h"
#include "interp_kernels.hpp"
#include "proposal.hpp"

struct brille_grid {
  // Every entry point that uses the handle holds `mu` for the whole call: the
  // error record, staging and work buffers, side stream and proposal state are
  // per handle, and the reference's ir_interpolate_at is const and re-entrant
  // (src/bz_trellis.hpp:153-155), so calls from several host threads on one
  // handle run one after the other and each gets its own results.
  mutable std::mutex mu;
  // bumped by every brille_grid_fill (its tables are replaced): a batch
  // refreshes its copy of a member's descriptors when this changes
  uint64_t fill_gen = 0;
  brille_amd::GridConst g{};
  brille_amd::DataConst d{};
  std::vector<void*> grid_allocs;
  std::vector<void*> data_allocs;
  // the permutation table brille_grid_sort built (freed by the next sort or fill)
  std::vector<void*> sort_allocs;
  uint32_t n_perms = 0;  // permutations of the installed table
  bool filled = false;
  uint64_t n_vertices = 0, n_cubes = 0, n_polys = 0, n_tets = 0;
  brille_amd::ErrRec* err = nullptr;
  std::vector<uint32_t> host_desc_poly_off;
  // staging buffers for host-memory calls
  void* stage = nullptr;
  size_t stage_bytes = 0;
  // work space of the binned pipeline (grown on demand, kept)
  void* bin_ws = nullptr;
  size_t bin_ws_bytes = 0;
  // side stream of the binned pipeline (preparation of the next pass) and its
  // events: call start, prepared[2], summed[2]
  hipStream_t bin_side = nullptr;
  hipEvent_t bin_ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint32_t n_cu = 0;  // compute units of the device (grid sizes of grid-stride kernels)
  // learned tetrahedron proposal: FP32 master parameters (flat: W1 b1 W2 b2
  // W3 b3), Adam moments and gradients; bf16 inference copies in `prop`
  brille_amd::ProposalConst prop{};
  std::vector<void*> prop_allocs;
  bool prop_trained = false;
  uint32_t prop_K = 0;
  uint64_t prop_n_params = 0;
  float* prop_params = nullptr;
  float* prop_m = nullptr;
  float* prop_v = nullptr;
  float* prop_grads = nullptr;
  uint64_t prop_step = 0;
  unsigned long long* prop_stats = nullptr;  // device: tried, verified
  unsigned long long last_tried = 0, last_hits = 0;
};

namespace brille_amd {
// the locks of several handles, taken in address order (no two callers can
// wait on each other)
class GridLocks {
 public:
  explicit GridLocks(std::vector<const brille_grid*> gs) {
    std::sort(gs.begin(), gs.end());
    gs.erase(std::unique(gs.begin(), gs.end()), gs.end());
    for (const brille_grid* g : gs)
      if (g) held_.emplace_back(g->mu);
  }

 private:
  std::vector<std::unique_lock<std::mutex>> held_;
};

// holds the handle's lock for the rest of the enclosing scope (no-op for null)
#define BRILLE_LOCK_GRID(G)                   \
  std::unique_lock<std::mutex> grid_lock_;    \
  if (G) grid_lock_ = std::unique_lock<std::mutex>((G)->mu)

int fail(int code, const std::string& msg);
const char* g_last_error_text();  // the message of the last failure on this thread
int build_fast_tables(const brille_grid_desc* desc, brille_grid* G);
void free_proposal(brille_grid* G);
void free_staging(brille_grid* G);
int check_rotation_modes(const brille_grid* G);
int check_payload_precision(const brille_grid* G, int precision);
// the fused tile kernel of one call (fast: tuned fold-in and sums; else the literal restatement)
void launch_tile(const brille_grid* G, const LaunchArgs& a, int precision, bool fast, hipStream_t s);
int payload_layout(const DataConst& d);
bool wide_rows(const DataConst& d);
// binned pipeline (binned.hip)
bool binned_applies(const brille_grid* G, const LaunchArgs& a, int kernel);
int launch_binned(brille_grid* G, const LaunchArgs& a, int precision, hipStream_t s, float* stage_ms);
// the pipeline on a given work space: passes of `chunk` points, `halves` = 2
// overlaps preparation and sums (grid side stream), 1 runs all on `s`
int run_binned_passes(brille_grid* G, const LaunchArgs& a, int precision, hipStream_t s, float* stage_ms, void* ws,
                      uint64_t chunk, int halves);
// bytes of one pass's buffers, and the largest pass whose buffers fit `bytes` (0: none)
uint64_t binned_pass_bytes(const brille_grid* G, uint64_t chunk);
uint64_t binned_pass_for(const brille_grid* G, uint64_t bytes);
void free_binned(brille_grid* G);
// clears n 32-bit words on stream s with a kernel (no memset node, so that a
// caller may capture the call in a graph)
void zero_words(void* p, uint32_t n, hipStream_t s);
// exclusive prefix sum of a[0, m) in place on stream s; part holds
// (m + 4095) / 4096 words; m at most 2^30
int exclusive_scan_u32(uint32_t* a, uint64_t m, uint32_t* part, hipStream_t s);
// the cube row table of the bucketed sums, built on first use (data_allocs)
int ensure_cube_rows(brille_grid* G, int precision, hipStream_t s);
// the first failing point of a call as the reference's error text
int report_point_error(const ErrRec& e, const double* Q, uint64_t n, bool host_mem);
// fold: FOLD_IR (ir_interpolate_at) or FOLD_FIRST_ZONE (interpolate_at); opts->no_move wins
int run_interpolate(brille_grid* G, const double* Q, uint64_t n, const brille_run_opts* opts, void* values_out,
                    void* vectors_out, const brille_aux_out* aux, hipStream_t s, brille_kernel_times* times,
                    uint32_t iters, int fold = FOLD_IR);
}  // namespace brille_amd
