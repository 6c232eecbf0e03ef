// First and irreducible Brillouin zones.
//
// Reference counterpart: BrillouinZone (src/bz.hpp:48-880, src/bz.cpp,
// src/bz_wedge.cpp).  The first zone is the Voronoi cell of the primitive
// reciprocal lattice, cut from a box by the bisecting planes of lattice
// vectors taken in order of increasing length (as voro_search does,
// src/bz.cpp:197-231).  The irreducible wedge is built differently from the
// reference's brute-force search: it is the Dirichlet domain of a generic
// seed point p0 under the point group, {q : q.(p0 - R^T p0) >= 0 for all R},
// a convex cone through the origin that tiles reciprocal space under the
// group.  Its facet normals are stored exactly like ir_wedge_normals, so the
// hot path's wedge test (src/bz.hpp:757-763) is unchanged.
#pragma once
#include <vector>

#include "lattice.hpp"
#include "polyhedron.hpp"

namespace brille_amd {

struct ApproxConfig {
  int digits = 1000;
  double direct = 1e-12;
  double reciprocal = 1e-12;
};

class BrillouinZone {
 public:
  Lattice outer;              // the lattice passed in (conventional)
  Lattice inner;              // primitive version used for the zone
  PointSymmetry psym;         // get_pointgroup_symmetry(): + inversion if time reversal
  bool time_reversal = false;
  bool has_inversion = false;
  bool transform_needed = false;
  bool wedge_searched = false;
  ConvexPolyhedron first;     // xyz
  std::vector<I3> first_taus; // per first-zone face: bisected lattice vector, primitive rlu
  ConvexPolyhedron ir;        // xyz
  std::vector<V3> wedge_normals_xyz;  // unit, pointing into the wedge
  V3 seed_xyz{0, 0, 0};
  bool has_seed = false;
  ApproxConfig approx;

  BrillouinZone() = default;
  BrillouinZone(const Lattice& lat, bool use_primitive = true, int search_length = 1,
                bool time_reversal = false, bool wedge_search = true, bool divide_primitive = true,
                ApproxConfig cfg = ApproxConfig());

  // coordinate conversions
  V3 xyz_from_outer(const V3& hkl) const { return mat_vec(outer.recip, hkl); }
  V3 outer_from_xyz(const V3& x) const { return mat_vec(inverse(outer.recip), x); }
  V3 prim_from_outer(const V3& hkl) const;  // (6 P^T hkl) / 6

  // per-face data of the first zone in conventional rlu (Faces::planes, face_points, face_normals)
  void first_planes_outer(std::vector<double>& a, std::vector<double>& b, std::vector<double>& c) const;
  std::vector<V3> first_normals_outer() const;
  std::vector<V3> first_points_outer() const;
  std::vector<V3> wedge_normals_outer() const;
  double first_volume() const { return first.volume(); }
  double ir_volume() const { return ir.volume(); }

 private:
  void voro_search(int extent);
  void build_wedge();
};

}  // namespace brille_amd
