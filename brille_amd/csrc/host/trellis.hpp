// Hybrid cube/tetrahedron trellis over the irreducible polyhedron.
//
// Reference counterpart: PolyTrellis::construct (src/trellis_poly_parallel.tpp:74-524).
// Knots are laid by repeated addition of step = extent / ceil(extent /
// cbrt(max_volume)) from the polyhedron's minimum (:86-102).  A node is a
// CubeNode when it lies wholly inside the polyhedron and does not contain
// Gamma (:235-237), a PolyNode when it intersects the polyhedron, otherwise
// null.  PolyNodes are the node box clipped by the polyhedron and split into
// tetrahedra by coning from an apex (Gamma when it lies inside, else a
// vertex) over the fan-triangulated faces that do not contain the apex; the
// reference uses TetGen without volume refinement for the same convex cells
// (src/triangulation_poly.hpp:158-163).
#pragma once
#include <array>
#include <vector>

#include "polyhedron.hpp"

namespace brille_amd {

struct TrellisTables {
  std::array<std::vector<double>, 3> knots;
  std::vector<uint8_t> node_type;
  std::vector<uint32_t> node_index;
  std::vector<uint32_t> cube_vertices;  // 8 per cube
  std::vector<uint32_t> poly_offset;    // n_polys + 1
  std::vector<uint32_t> poly_node;      // linear node index of each poly
  std::vector<uint32_t> tet_vertices;   // 4 per tet
  std::vector<double> tet_circum;       // 4 per tet
  std::vector<double> tet_volume;
  std::vector<double> vertices;         // 3 per vertex, xyz
  uint64_t n_vertices() const { return vertices.size() / 3; }
  uint64_t n_nodes() const { return node_type.size(); }
  uint64_t n_cubes() const { return cube_vertices.size() / 8; }
  uint64_t n_polys() const { return poly_offset.empty() ? 0 : poly_offset.size() - 1; }
  uint64_t n_tets() const { return tet_volume.size(); }
  double total_node_volume() const;
};

TrellisTables build_trellis(const ConvexPolyhedron& poly, double max_volume, bool always_triangulate);

}  // namespace brille_amd
