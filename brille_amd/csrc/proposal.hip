// Learned tetrahedron proposal (no reference counterpart).
#include <hip/hip_runtime.h>

#include "grid_state.hpp"

using namespace brille_amd;

extern "C" {

int brille_proposal_train(brille_grid* G, const brille_train_opts* opts, brille_train_report* report) {
  (void)G;
  (void)opts;
  (void)report;
  return fail(BRILLE_ERR_UNSUPPORTED, "the learned proposal is not built yet");
}

int brille_proposal_loss_grad(brille_grid* G, const float* x, const int32_t* labels, uint32_t n, float* loss,
                              float* grads, uint32_t* n_params) {
  (void)G; (void)x; (void)labels; (void)n; (void)loss; (void)grads; (void)n_params;
  return fail(BRILLE_ERR_UNSUPPORTED, "the learned proposal is not built yet");
}

int brille_proposal_get_params(const brille_grid* G, float* params, uint32_t* n_params) {
  (void)G; (void)params; (void)n_params;
  return fail(BRILLE_ERR_UNSUPPORTED, "the learned proposal is not built yet");
}

int brille_proposal_set_params(brille_grid* G, const float* params, uint32_t n_params) {
  (void)G; (void)params; (void)n_params;
  return fail(BRILLE_ERR_UNSUPPORTED, "the learned proposal is not built yet");
}

}  // extern "C"
