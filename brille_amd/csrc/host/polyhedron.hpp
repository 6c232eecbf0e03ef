// Convex polyhedra built by successive half-space clipping (Cartesian xyz).
//
// Plays the role of the reference's polyhedron::Poly for the shapes the path
// needs: the first Brillouin zone, the irreducible wedge polyhedron and each
// trellis node intersected with it (src/polyhedron_flex.hpp,
// src/polyhedron_faces.hpp).  Faces are stored with their vertices ordered
// counter-clockwise when seen from outside, so cross(b-a, c-b) of the first
// three vertices is the outward normal (src/geometry.hpp:249-252).
#pragma once
#include <vector>

#include "linalg.hpp"

namespace brille_amd {

struct Plane {
  V3 n{0, 0, 0};  // outward normal
  double d = 0;   // inside: n.x <= d
  int tag = -1;   // caller's label (e.g. the lattice vector a zone face bisects)
};

class ConvexPolyhedron {
 public:
  std::vector<V3> verts;
  std::vector<std::vector<uint32_t>> faces;
  std::vector<Plane> planes;  // one per face

  static ConvexPolyhedron box(const V3& lo, const V3& hi);
  // keep the part with n.x <= d; `tol` is an absolute distance tolerance
  ConvexPolyhedron clip(const Plane& p, double tol) const;
  bool empty() const { return faces.size() < 4; }
  double volume() const;
  V3 centroid() const;  // vertex average
  double extent() const;  // largest |coordinate|
  // inside or on every face plane within tol
  bool contains(const V3& x, double tol) const;
  // signed distance beyond the most violated face plane (<= 0 inside)
  double max_violation(const V3& x) const;
};

}  // namespace brille_amd
