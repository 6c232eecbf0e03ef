// CPU harness (test infrastructure): runs the kernels' fold-in and location
// routines on the host, the exact and the fast variants side by side, so the
// fast decisions can be checked against the exact ones without a GPU.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstring>
#include <limits>

#include "../../brille_amd/csrc/exact_path.hpp"
#include "../../brille_amd/csrc/fast_fold.hpp"
#include "../../brille_amd/csrc/fast_tables.hpp"

using namespace brille_amd;

static void tols(double Ttol, int tol, double* rel, double* ab) {
  double Trel = std::numeric_limits<double>::epsilon(), Tabs = 5.0 / 1000000000000000.0;
  Trel *= (double)tol * 10000.0;
  if (Ttol > Trel) Trel = Ttol;
  if (Ttol > Tabs) Tabs = Ttol;
  *rel = Trel;
  *ab = Tabs;
}

static void host_grid(const brille_grid_desc* d, const FastTables& ft, GridConst& g) {
  std::memset(&g, 0, sizeof(g));
  std::memcpy(g.outer_to_xyz, d->outer_to_xyz, sizeof(g.outer_to_xyz));
  std::memcpy(g.outer_Gr, d->outer_metric_recip, sizeof(g.outer_Gr));
  std::memcpy(g.outer_Gd, d->outer_metric_real, sizeof(g.outer_Gd));
  g.outer_scale = d->outer_cross_scale;
  std::memcpy(g.prim_Gr, d->prim_metric_recip, sizeof(g.prim_Gr));
  std::memcpy(g.prim_Gd, d->prim_metric_real, sizeof(g.prim_Gd));
  g.prim_scale = d->prim_cross_scale;
  std::memcpy(g.P6t, d->to_primitive_6Pt, sizeof(g.P6t));
  std::memcpy(g.invPt, d->from_primitive_invPt, sizeof(g.invPt));
  g.transform_needed = d->transform_needed;
  g.no_ir_mirroring = d->no_ir_mirroring;
  g.has_seed = ft.has_orbit ? 1 : 0;
  g.F = d->n_faces;
  g.W = d->n_wedge;
  g.G = d->n_ops;
  g.identity = d->identity_index;
  tols(d->float_tolerance, d->approx_digits, &g.tol_bz_rel, &g.tol_bz_abs);
  tols(0.0, 1, &g.tol_def_rel, &g.tol_def_abs);
  g.prim_pa = d->prim_plane_a; g.prim_pb = d->prim_plane_b; g.prim_pc = d->prim_plane_c;
  g.prim_n = d->prim_normals; g.prim_taus = d->prim_taus; g.prim_tau_lens = d->prim_tau_lens;
  g.outer_pa = d->outer_plane_a; g.outer_pb = d->outer_plane_b; g.outer_pc = d->outer_plane_c;
  g.wedge_n = d->wedge_normals; g.rot = d->rotations; g.inv_index = d->inverse_index;
  g.prim_aff = ft.prim_aff.data(); g.outer_aff = ft.outer_aff.data();
  g.orbit = ft.orbit.empty() ? nullptr : ft.orbit.data();
  g.aff_margin = ft.aff_margin; g.orbit_margin = ft.orbit_margin;
  for (int a = 0; a < 3; ++a) {
    g.nk[a] = d->n_knots[a];
    g.knots[a] = d->knots[a];
    if (d->n_knots[a] >= 2) {
      g.knot_lo[a] = d->knots[a][0];
      double step = (d->knots[a][d->n_knots[a] - 1] - d->knots[a][0]) / (double)(d->n_knots[a] - 1);
      g.knot_inv_step[a] = step > 0 ? 1.0 / step : 0.0;
    }
  }
  g.node_type = d->node_type; g.node_index = d->node_index; g.cube_v = d->cube_vertices;
  g.poly_off = d->poly_tet_offset; g.tet_v = d->tet_vertices; g.tet_c = d->tet_circumsphere;
  g.tet_vol = d->tet_volume; g.verts = d->vertices; g.n_nodes = d->n_nodes;
}

extern "C" int fold_check(const brille_grid_desc* d, const double* Q, uint64_t n, double* q, int32_t* tau,
                          uint32_t* rot, uint32_t* invrot, uint8_t* st, uint64_t* nodes, uint64_t* counts) {
  FastTables ft = compute_fast_tables(d);
  GridConst g;
  host_grid(d, ft, g);
  const Tol tbz{g.tol_bz_rel, g.tol_bz_abs};
  const Tol tdef{g.tol_def_rel, g.tol_def_abs};
  uint64_t mismatches = 0, undecided_inside = 0;
  for (uint64_t i = 0; i < n; ++i) {
    // exact (slot 0) and fast (slot 1)
    for (int fast = 0; fast < 2; ++fast) {
      double qq[3];
      int32_t tt[3];
      uint32_t r = g.identity, ir = g.identity;
      int s = fast ? fast_moveinto(g, Q + 3 * i, qq, tt, tbz) : moveinto_exact(g, Q + 3 * i, qq, tt, tbz);
      if (s == BRILLE_PT_OK)
        s = fast ? fast_wedge_search(g, qq, &r, &ir, tbz, tdef) : wedge_search_exact(g, qq, &r, &ir, tbz, tdef);
      const uint64_t k = 2 * i + fast;
      std::memcpy(q + 3 * k, qq, sizeof(qq));
      std::memcpy(tau + 3 * k, tt, sizeof(tt));
      rot[k] = r;
      invrot[k] = ir;
      st[k] = (uint8_t)s;
      uint64_t node = ~0ull;
      if (s == BRILLE_PT_OK && g.n_nodes) {
        double x[3];
        mat_vec(x, g.outer_to_xyz, qq);
        bool ok = fast ? node_subscript<true>(g, x, &node, tdef) : node_subscript<false>(g, x, &node, tdef);
        if (!ok) node = ~1ull;
      }
      nodes[k] = node;
    }
    if (std::memcmp(q + 6 * i, q + 6 * i + 3, 3 * sizeof(double)) || std::memcmp(tau + 6 * i, tau + 6 * i + 3, 12) ||
        rot[2 * i] != rot[2 * i + 1] || invrot[2 * i] != invrot[2 * i + 1] || st[2 * i] != st[2 * i + 1] ||
        nodes[2 * i] != nodes[2 * i + 1])
      ++mismatches;
    double qd[3] = {Q[3 * i], Q[3 * i + 1], Q[3 * i + 2]};
    if (fast_inside(g.F, g.outer_aff, qd, g.aff_margin) == 0) ++undecided_inside;
  }
  counts[0] = mismatches;
  counts[1] = undecided_inside;
  return 0;
}
