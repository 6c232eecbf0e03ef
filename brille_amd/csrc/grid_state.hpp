// Host-side state behind the opaque brille_grid / brille_batch handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/brille_amd.h"
#include "interp_kernels.hpp"

struct brille_grid {
  brille_amd::GridConst g{};
  brille_amd::DataConst d{};
  std::vector<void*> grid_allocs;
  std::vector<void*> data_allocs;
  bool filled = false;
  uint64_t n_vertices = 0, n_polys = 0, n_tets = 0;
  brille_amd::ErrRec* err = nullptr;
  std::vector<uint32_t> host_desc_poly_off;
  // staging buffers for host-memory calls
  void* stage = nullptr;
  size_t stage_bytes = 0;
  // learned tetrahedron proposal
  brille_amd::ProposalConst prop{};
  std::vector<void*> prop_allocs;
  bool prop_trained = false;
  unsigned long long* prop_stats = nullptr;  // device: tried, hits
  unsigned long long last_tried = 0, last_hits = 0;
};

namespace brille_amd {
int fail(int code, const std::string& msg);
int build_fast_tables(const brille_grid_desc* desc, brille_grid* G);
void free_proposal(brille_grid* G);
void free_staging(brille_grid* G);
int check_rotation_modes(const brille_grid* G);
// the first failing point of a call as the reference's error text
int report_point_error(const ErrRec& e, const double* Q, uint64_t n, bool host_mem);
int run_interpolate(brille_grid* G, const double* Q, uint64_t n, const brille_run_opts* opts, void* values_out,
                    void* vectors_out, const brille_aux_out* aux, hipStream_t s, brille_kernel_times* times,
                    uint32_t iters);
}  // namespace brille_amd
