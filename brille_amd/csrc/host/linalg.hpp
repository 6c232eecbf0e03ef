// Small fixed-size linear algebra used by the host-side constructors.
// Matrices are 3x3 row-major std::array<T,9>; vectors std::array<T,3>.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>

namespace brille_amd {

using V3 = std::array<double, 3>;
using M3 = std::array<double, 9>;
using I3 = std::array<int32_t, 3>;
using IM3 = std::array<int32_t, 9>;

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kPi = 3.14159265358979323846264338327950;

template <class A, class B>
inline auto mat_mul(const std::array<A, 9>& a, const std::array<B, 9>& b) {
  using T = decltype(a[0] * b[0]);
  std::array<T, 9> c{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = T(0);
      for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
      c[3 * i + j] = s;
    }
  return c;
}

template <class A, class B>
inline auto mat_vec(const std::array<A, 9>& a, const std::array<B, 3>& v) {
  using T = decltype(a[0] * v[0]);
  std::array<T, 3> c{};
  for (int i = 0; i < 3; ++i) {
    T s = T(0);
    for (int k = 0; k < 3; ++k) s += a[3 * i + k] * v[k];
    c[i] = s;
  }
  return c;
}

template <class T>
inline std::array<T, 9> transpose(const std::array<T, 9>& a) {
  return {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
}

template <class T>
inline T det3(const std::array<T, 9>& m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

inline M3 inverse(const M3& m) {
  double d = det3(m);
  M3 r{};
  r[0] = (m[4] * m[8] - m[5] * m[7]) / d;
  r[1] = (m[2] * m[7] - m[1] * m[8]) / d;
  r[2] = (m[1] * m[5] - m[2] * m[4]) / d;
  r[3] = (m[5] * m[6] - m[3] * m[8]) / d;
  r[4] = (m[0] * m[8] - m[2] * m[6]) / d;
  r[5] = (m[2] * m[3] - m[0] * m[5]) / d;
  r[6] = (m[3] * m[7] - m[4] * m[6]) / d;
  r[7] = (m[1] * m[6] - m[0] * m[7]) / d;
  r[8] = (m[0] * m[4] - m[1] * m[3]) / d;
  return r;
}

inline IM3 inverse_int(const IM3& m) {
  // integer matrices of point-group operations have determinant +-1
  int32_t d = det3(m);
  IM3 r{};
  r[0] = (m[4] * m[8] - m[5] * m[7]) * d;
  r[1] = (m[2] * m[7] - m[1] * m[8]) * d;
  r[2] = (m[1] * m[5] - m[2] * m[4]) * d;
  r[3] = (m[5] * m[6] - m[3] * m[8]) * d;
  r[4] = (m[0] * m[8] - m[2] * m[6]) * d;
  r[5] = (m[2] * m[3] - m[0] * m[5]) * d;
  r[6] = (m[3] * m[7] - m[4] * m[6]) * d;
  r[7] = (m[1] * m[6] - m[0] * m[7]) * d;
  r[8] = (m[0] * m[4] - m[1] * m[3]) * d;
  return r;
}

inline M3 to_double(const IM3& m) {
  M3 r{};
  for (int i = 0; i < 9; ++i) r[i] = static_cast<double>(m[i]);
  return r;
}

inline V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 add(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 scale(const V3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }

inline bool is_identity(const IM3& m) {
  return m == IM3{1, 0, 0, 0, 1, 0, 0, 0, 1};
}

}  // namespace brille_amd
