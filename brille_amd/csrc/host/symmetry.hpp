// Space-group and point-group bookkeeping for the host-side constructors.
//
// Reference counterparts: Motion (src/motion.hpp:27), Symmetry
// (src/symmetry.hpp:69), HallSymbol (src/hall_symbol.hpp:181), PointSymmetry
// (src/pointsymmetry.hpp:44-48, src/pointsymmetry.cpp:131-255), Basis
// (src/basis.hpp:35-200).  Space groups are generated from Hall symbols
// (S. R. Hall, Acta Cryst. A37, 517 (1981)) by a parser written here.
#pragma once
#include <string>
#include <vector>

#include "linalg.hpp"

namespace brille_amd {

enum class Bravais { _ = 0, P = 1, A = 2, B = 3, C = 4, I = 5, F = 6, R = 7 };

struct Motion {
  IM3 W{1, 0, 0, 0, 1, 0, 0, 0, 1};
  V3 w{0, 0, 0};  // translation in [0, 1)
  V3 move_point(const V3& p) const {
    V3 o = mat_vec(to_double(W), p);
    for (int i = 0; i < 3; ++i) o[i] += w[i];
    return o;
  }
  Motion inverse() const;
  Motion operator*(const Motion& o) const;  // (this) after (o)
  bool equals(const Motion& o, double tol = 1e-10) const;
};

class Symmetry {
 public:
  std::vector<Motion> ops;
  Symmetry() = default;
  explicit Symmetry(std::vector<Motion> m) : ops(std::move(m)) {}
  // all matrices W (n,9) and translations w (n,3)
  static Symmetry from_matrices(const std::vector<IM3>& W, const std::vector<V3>& w);
  Symmetry generate() const;  // closure under composition, translations mod 1
  size_t size() const { return ops.size(); }
  Bravais centring() const;
  size_t find_matrix_index(const IM3& R) const;
  bool has_space_inversion() const;
  Symmetry add_space_inversion() const;
};

// Parse a Hall symbol into its generators (centring translations included).
Symmetry hall_generators(const std::string& hall, Bravais* bravais = nullptr);
// Full space group from a Hall symbol, an ITA number ("194") or one of the
// Hermann-Mauguin names in the built-in table.
Symmetry spacegroup_from_string(const std::string& s, Bravais* bravais = nullptr);
// Default-setting Hall symbol of an ITA number, "" when not tabulated.
std::string hall_symbol_of(int ita_number);
// ITA number named by a space-group string (number, Hall symbol of a default
// setting, Hermann-Mauguin symbol); 0 when the string names another setting
// or nothing.
int spacegroup_number_of(const std::string& symbol);

int rotation_order(const IM3& R);
int isometry_value(const IM3& R);  // +-order, negative for improper

class PointSymmetry {
 public:
  std::vector<IM3> R;
  PointSymmetry() = default;
  explicit PointSymmetry(std::vector<IM3> r) : R(std::move(r)) {}
  static PointSymmetry from_spacegroup(const Symmetry& s);  // unique rotations, sorted
  size_t size() const { return R.size(); }
  size_t find_index(const IM3& a) const;
  size_t find_identity_index() const;
  size_t get_inverse_index(size_t i) const;
  bool has_space_inversion() const;
  PointSymmetry add_space_inversion() const;
  void sort_by_order();  // increasing rotation order; stable
  std::vector<uint32_t> inverse_table() const;
};

class Basis {
 public:
  std::vector<V3> positions;
  std::vector<uint32_t> types;
  Basis() = default;
  Basis(std::vector<V3> p, std::vector<uint32_t> t) : positions(std::move(p)), types(std::move(t)) {}
  size_t size() const { return positions.size(); }
  // src/basis.hpp:140-190: equivalent atom index of k after the operation
  bool equivalent_after_operation(size_t k, const Motion& op, double e_tol, int n_tol, uint32_t* l) const;
};

}  // namespace brille_amd
