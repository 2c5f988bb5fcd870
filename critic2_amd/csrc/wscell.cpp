// Wigner-Seitz cell of a lattice: its facet vectors (the Voronoi-relevant
// lattice vectors) and facet areas, what grid3%vec / grid3%area hold for the
// grid lattice (wscell, tools@proc.f90:679-800, there through qhull).
//
// The facet of a lattice vector v is the part of the plane x.v = |v|^2/2
// where x.u <= |u|^2/2 holds for every other lattice vector u.  Each
// candidate v = i a + j b + k c (|i|,|j|,|k| <= 3) gets its facet by
// clipping a large square in that plane with the half-planes of all the
// other candidates; v is relevant when the facet keeps a positive area.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace c2 {

namespace {

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct P2 {
  double s, t;
};

// Keeps the part of a convex polygon with a s + b t <= c.
std::vector<P2> clip(const std::vector<P2> &poly, double a, double b, double c) {
  std::vector<P2> out;
  const size_t m = poly.size();
  for (size_t q = 0; q < m; q++) {
    const P2 p = poly[q], r = poly[(q + 1) % m];
    const double fp = a * p.s + b * p.t - c, fr = a * r.s + b * r.t - c;
    if (fp <= 0) out.push_back(p);
    if ((fp < 0 && fr > 0) || (fp > 0 && fr < 0)) {
      const double u = fp / (fp - fr);
      out.push_back({p.s + u * (r.s - p.s), p.t + u * (r.t - p.t)});
    }
  }
  return out;
}

}  // namespace

// m: lattice vectors as columns (column-major).  Returns the facet count;
// vec: 3 integers per facet (lattice coordinates), area: its area.
int ws_facets(const double m[9], int vec[3 * 14], double area[14]) {
  const V3 a{m[0], m[1], m[2]}, b{m[3], m[4], m[5]}, c{m[6], m[7], m[8]};
  const double vol = std::fabs(dot(a, cross(b, c)));
  if (!(vol > 0)) fail(C2GPU_ERR_ARG, "Wigner-Seitz cell: singular lattice");
  struct Cand {
    int i, j, k;
    V3 v;
  };
  std::vector<Cand> cand;
  const int R = 3;
  for (int i = -R; i <= R; i++)
    for (int j = -R; j <= R; j++)
      for (int k = -R; k <= R; k++)
        if (i || j || k) cand.push_back({i, j, k, (double)i * a + (double)j * b + (double)k * c});
  const double big = 4.0 * (std::sqrt(dot(a, a)) + std::sqrt(dot(b, b)) + std::sqrt(dot(c, c)));
  int nf = 0;
  double vsum = 0.0;  // the cell volume from its facets: sum area * |v| / 6
  for (const Cand &cv : cand) {
    const V3 v = cv.v;
    const double vv = dot(v, v);
    // orthonormal basis of the plane
    V3 e1 = std::fabs(v.x) < 0.9 * std::sqrt(vv) ? cross(v, V3{1, 0, 0}) : cross(v, V3{0, 1, 0});
    e1 = (1.0 / std::sqrt(dot(e1, e1))) * e1;
    V3 e2 = cross(v, e1);
    e2 = (1.0 / std::sqrt(dot(e2, e2))) * e2;
    const V3 cen = 0.5 * v;
    std::vector<P2> poly{{-big, -big}, {big, -big}, {big, big}, {-big, big}};
    for (const Cand &cu : cand) {
      if (&cu == &cv) continue;
      const V3 u = cu.v;
      poly = clip(poly, dot(e1, u), dot(e2, u), 0.5 * dot(u, u) - dot(cen, u));
      if (poly.size() < 3) break;
    }
    if (poly.size() < 3) continue;
    double ar = 0.0;
    for (size_t q = 0; q < poly.size(); q++) {
      const P2 p = poly[q], r = poly[(q + 1) % poly.size()];
      ar += p.s * r.t - r.s * p.t;
    }
    ar = 0.5 * std::fabs(ar);
    if (ar <= 1e-10 * std::cbrt(vol * vol)) continue;
    if (nf >= 14) fail(C2GPU_ERR_ARG, "Wigner-Seitz cell: more than 14 facets");
    vec[3 * nf] = cv.i;
    vec[3 * nf + 1] = cv.j;
    vec[3 * nf + 2] = cv.k;
    area[nf] = ar;
    vsum += ar * std::sqrt(vv) / 6.0;
    nf++;
  }
  if (std::fabs(vsum - vol) > 1e-8 * vol)
    fail(C2GPU_ERR_ARG, "Wigner-Seitz cell: facets do not close (lattice too far from reduced)");
  return nf;
}

}  // namespace c2
