#include "polyhedron.hpp"

#include <algorithm>
#include <cmath>

namespace brille_amd {

namespace {

struct Face {
  std::vector<V3> pts;
  Plane plane;
};

void order_ccw(std::vector<V3>& pts, const V3& n) {
  if (pts.size() < 3) return;
  V3 c{0, 0, 0};
  for (const auto& p : pts) c = add(c, p);
  c = scale(c, 1.0 / static_cast<double>(pts.size()));
  // pick the point farthest from the centre as the reference direction
  size_t far = 0;
  double best = -1;
  for (size_t i = 0; i < pts.size(); ++i) {
    double d = norm(sub(pts[i], c));
    if (d > best) { best = d; far = i; }
  }
  V3 u = sub(pts[far], c);
  u = scale(u, 1.0 / norm(u));
  V3 nn = scale(n, 1.0 / norm(n));
  V3 v = cross(nn, u);
  std::vector<std::pair<double, V3>> ang;
  ang.reserve(pts.size());
  for (const auto& p : pts) {
    V3 r = sub(p, c);
    ang.emplace_back(std::atan2(dot(r, v), dot(r, u)), p);
  }
  std::sort(ang.begin(), ang.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  for (size_t i = 0; i < pts.size(); ++i) pts[i] = ang[i].second;
}

void dedupe_points(std::vector<V3>& pts, double tol) {
  std::vector<V3> out;
  for (const auto& p : pts) {
    bool dup = false;
    for (const auto& q : out)
      if (std::abs(p[0] - q[0]) <= tol && std::abs(p[1] - q[1]) <= tol && std::abs(p[2] - q[2]) <= tol) {
        dup = true;
        break;
      }
    if (!dup) out.push_back(p);
  }
  pts.swap(out);
}

ConvexPolyhedron assemble(const std::vector<Face>& faces, double tol) {
  ConvexPolyhedron P;
  for (const auto& f : faces) {
    if (f.pts.size() < 3) continue;
    std::vector<uint32_t> idx;
    for (const auto& p : f.pts) {
      uint32_t found = static_cast<uint32_t>(P.verts.size());
      for (uint32_t i = 0; i < P.verts.size(); ++i) {
        const V3& q = P.verts[i];
        if (std::abs(p[0] - q[0]) <= tol && std::abs(p[1] - q[1]) <= tol && std::abs(p[2] - q[2]) <= tol) {
          found = i;
          break;
        }
      }
      if (found == P.verts.size()) P.verts.push_back(p);
      if (std::find(idx.begin(), idx.end(), found) == idx.end()) idx.push_back(found);
    }
    if (idx.size() >= 3) {
      P.faces.push_back(idx);
      P.planes.push_back(f.plane);
    }
  }
  return P;
}

}  // namespace

ConvexPolyhedron ConvexPolyhedron::box(const V3& lo, const V3& hi) {
  std::vector<Face> faces(6);
  V3 c[8];
  for (int i = 0; i < 8; ++i) c[i] = {(i & 1) ? hi[0] : lo[0], (i & 2) ? hi[1] : lo[1], (i & 4) ? hi[2] : lo[2]};
  auto mk = [&](int f, V3 n, double d, std::initializer_list<int> ids) {
    faces[f].plane.n = n;
    faces[f].plane.d = d;
    for (int i : ids) faces[f].pts.push_back(c[i]);
    order_ccw(faces[f].pts, n);
  };
  mk(0, {-1, 0, 0}, -lo[0], {0, 2, 4, 6});
  mk(1, {1, 0, 0}, hi[0], {1, 3, 5, 7});
  mk(2, {0, -1, 0}, -lo[1], {0, 1, 4, 5});
  mk(3, {0, 1, 0}, hi[1], {2, 3, 6, 7});
  mk(4, {0, 0, -1}, -lo[2], {0, 1, 2, 3});
  mk(5, {0, 0, 1}, hi[2], {4, 5, 6, 7});
  double s = std::max({std::abs(lo[0]), std::abs(lo[1]), std::abs(lo[2]), std::abs(hi[0]), std::abs(hi[1]), std::abs(hi[2])});
  return assemble(faces, 1e-14 * (s > 0 ? s : 1.0));
}

ConvexPolyhedron ConvexPolyhedron::clip(const Plane& p, double tol) const {
  // classify vertices
  bool any_out = false, all_out = true;
  std::vector<double> s(verts.size());
  double nn = norm(p.n);
  for (size_t i = 0; i < verts.size(); ++i) {
    s[i] = (dot(p.n, verts[i]) - p.d) / nn;
    if (s[i] > tol) any_out = true;
    if (s[i] <= tol) all_out = false;
  }
  if (!any_out) return *this;
  if (all_out) return ConvexPolyhedron{};
  std::vector<Face> out;
  std::vector<V3> cap;
  for (size_t f = 0; f < faces.size(); ++f) {
    const auto& face = faces[f];
    Face nf;
    nf.plane = planes[f];
    const size_t m = face.size();
    for (size_t k = 0; k < m; ++k) {
      uint32_t a = face[k], b = face[(k + 1) % m];
      double sa = s[a], sb = s[b];
      const V3& A = verts[a];
      const V3& B = verts[b];
      bool ina = sa <= tol, inb = sb <= tol;
      if (ina) {
        nf.pts.push_back(A);
        if (sa >= -tol) cap.push_back(A);
      }
      if ((sa < -tol && sb > tol) || (sa > tol && sb < -tol)) {
        double t = sa / (sa - sb);
        V3 X = add(A, scale(sub(B, A), t));
        nf.pts.push_back(X);
        cap.push_back(X);
      }
      (void)inb;
    }
    dedupe_points(nf.pts, tol);
    if (nf.pts.size() >= 3) out.push_back(nf);
  }
  dedupe_points(cap, tol);
  if (cap.size() >= 3) {
    Face cf;
    cf.plane = p;
    cf.pts = cap;
    order_ccw(cf.pts, p.n);
    out.push_back(cf);
  }
  ConvexPolyhedron r = assemble(out, tol);
  if (r.faces.size() < 4) return ConvexPolyhedron{};
  return r;
}

double ConvexPolyhedron::volume() const {
  if (faces.size() < 4) return 0.0;
  V3 c = centroid();
  double v = 0;
  for (const auto& f : faces) {
    for (size_t i = 1; i + 1 < f.size(); ++i) {
      V3 a = sub(verts[f[0]], c), b = sub(verts[f[i]], c), d = sub(verts[f[i + 1]], c);
      v += dot(a, cross(b, d));
    }
  }
  return v / 6.0;
}

V3 ConvexPolyhedron::centroid() const {
  V3 c{0, 0, 0};
  if (verts.empty()) return c;
  for (const auto& p : verts) c = add(c, p);
  return scale(c, 1.0 / static_cast<double>(verts.size()));
}

double ConvexPolyhedron::extent() const {
  double s = 0;
  for (const auto& p : verts)
    for (double x : p) s = std::max(s, std::abs(x));
  return s;
}

bool ConvexPolyhedron::contains(const V3& x, double tol) const { return max_violation(x) <= tol; }

double ConvexPolyhedron::max_violation(const V3& x) const {
  double m = -1e300;
  for (const auto& p : planes) m = std::max(m, (dot(p.n, x) - p.d) / norm(p.n));
  return m;
}

}  // namespace brille_amd
