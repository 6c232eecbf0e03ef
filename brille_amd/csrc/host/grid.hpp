// BZTrellisQdc: a BrillouinZone, its trellis and the interpolation payload,
// with the device handle that runs ir_interpolate_at.
//
// Reference counterpart: BrillouinZoneTrellis3<double, std::complex<double>, double>
// (src/bz_trellis.hpp:39-233) with DualInterpolator data (src/interpolatordual.hpp)
// and the GammaTable built per call (src/bz_trellis.hpp:182-186, src/phonon.hpp:106-167).
#pragma once
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/brille_amd.h"
#include "bz.hpp"
#include "trellis.hpp"

namespace brille_amd {

// GammaTable  src/phonon.hpp:91-194
struct GammaTableHost {
  uint32_t n_atoms = 0, n_ops = 0;
  std::vector<uint32_t> F0, vindex;
  std::vector<double> vectors;        // n_vectors x 3
  std::vector<double> cart_rotations; // n_ops x 9
  void construct(const Lattice& lat, const PointSymmetry& ps, bool time_reversal, double e_tol, int n_tol);
};

// Flattened zone tables (the brille_grid_desc arrays that are not trellis).
struct ZoneTables {
  std::vector<double> prim_a, prim_b, prim_c, prim_n, prim_tau_lens, outer_a, outer_b, outer_c, wedge;
  std::vector<int32_t> prim_taus, rotations;
  std::vector<uint32_t> inverse;
  void build(const BrillouinZone& bz);
  // fill every zone field of desc (pointers into this object)
  void describe(const BrillouinZone& bz, brille_grid_desc& d) const;
};

struct DeviceGrid {
  brille_grid* h = nullptr;
  ~DeviceGrid();
};

std::string last_error_text();
void check_status(int rc);

struct Payload {
  std::vector<double> data;     // real, or interleaved complex
  std::vector<size_t> shape;    // (V, B, ...)
  uint32_t elements[3] = {0, 0, 0};
  int rot = BRILLE_ROT_VECTOR;
  int lu = BRILLE_LU_REAL_LATTICE;
  uint32_t branches = 0;
  // mode-assignment cost (Interpolator::set_cost_info and _costmult,
  // src/interpolator.hpp:246-305); the reference's defaults unless a
  // weighted fill gave others
  brille_cost_info cost{0, 0, {1.0, 1.0, 1.0}};
  uint32_t span() const { return elements[0] + elements[1] + elements[2]; }
};

// Interpolator::check_elements / branches  src/interpolator.hpp:309-332, 547-584
void parse_payload(Payload& p, const std::vector<size_t>& shape, const std::vector<int>& elements, bool complex_data);

class BZTrellisQdc {
 public:
  BrillouinZone bz;
  TrellisTables trellis;
  ApproxConfig approx;
  Payload values, vectors;
  bool filled = false;
  GammaTableHost gamma;
  ZoneTables zone;
  std::shared_ptr<DeviceGrid> device;
  bool device_filled = false;
  // fill(..., sort=True): the device table is sorted right after the device fill
  bool sort_requested = false;
  // guards the lazy device create + fill and the host payload it reads:
  // callers release the GIL around interpolation, so several threads can
  // reach handle() at once
  std::shared_ptr<std::mutex> host_mu = std::make_shared<std::mutex>();

  BZTrellisQdc(const BrillouinZone& bz, double node_volume_fraction, bool always_triangulate, ApproxConfig cfg);
  void fill(Payload v, Payload c, bool sort = false);
  brille_grid_desc describe() const;
  brille_grid* handle();  // create + fill on first use
  uint64_t bytes_per_point(int precision = BRILLE_FP64) const;
  // Q: n x 3 conventional rlu; outputs sized n * B * span (real) and n * B * span * 2 (complex)
  // first_zone: BrillouinZoneTrellis3::interpolate_at (moveinto only, no rotation)
  void ir_interpolate_at(const double* Q, uint64_t n, bool no_move, int precision, int kernel, void* vals, void* vecs,
                         const brille_aux_out* aux = nullptr, bool first_zone = false);
};

// zone-only device handle for BrillouinZone.moveinto / ir_moveinto
std::shared_ptr<DeviceGrid> zone_handle(const BrillouinZone& bz, ZoneTables& store);

}  // namespace brille_amd
