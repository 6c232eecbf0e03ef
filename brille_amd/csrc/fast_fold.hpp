// Fast stage (a): the same decisions as the exact fold-in, taken on cheap
// bounds first and on the exact reference arithmetic only when a bound is
// too close to call.
#pragma once
#include "exact_path.hpp"

namespace brille_amd {

__device__ inline int fast_moveinto(const GridConst& g, const double* Q, double* q, int32_t* tau, Tol tbz) {
  return moveinto_exact(g, Q, q, tau, tbz);
}

__device__ inline int fast_wedge_search(const GridConst& g, double* q, uint32_t* rot, uint32_t* invrot, Tol tbz, Tol tdef) {
  return wedge_search_exact(g, q, rot, invrot, tbz, tdef);
}

}  // namespace brille_amd
