// Cone-segment grid of one tree level, shared by the host tree build and the
// device relevant-cone build: the centre of a box from its Morton code
// (Level::GetBoxCenter, IFGF_Source/Level.h:291-298), the non-relevant cone
// holding a point (Level::Cart2Nonrelcone, Level.h:310-323, with
// Functions::Atan2, PolarCoordinates.h:8-27) and the Cartesian position of a
// cone's interpolation node (Level::Cheb2Cart, Level.h:347-375).
//
// Every function runs the same IEEE operations on the host and on the GPU:
// multiply-add contraction is switched off in each body, division and square
// root are correctly rounded on both sides. atan, cos and sin come from each
// side's math library and may differ in the last bit, which moves a cone index
// only for a point within a few ulps of a cone boundary; grid_nonrel_near
// reports such points so that a device builder can leave them to the host.
#pragma once

#include <cmath>
#include <cstdint>

#include "zmath.h"

namespace ifgf {

// Inverse of the Morton interleave (Level.h:172-237): the x bits of a code.
IFGF_HD inline uint64_t morton_merge2(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x ^ (x >> 1)) & 0x3333333333333333ull;
  x = (x ^ (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x ^ (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x ^ (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x ^ (x >> 16)) & 0x00000000FFFFFFFFull;
  return x;
}

// Morton interleave of box (px, py) at a level (Level.h:172-237): x in the
// low bit, level + 1 bits per coordinate.
IFGF_HD inline uint64_t morton_split2(uint64_t v, int level) {
  uint64_t x = v & ((uint64_t(1) << (level + 1)) - 1);
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
IFGF_HD inline long long morton_code(long long px, long long py, int level) {
  return (long long)(morton_split2((uint64_t)px, level) | (morton_split2((uint64_t)py, level) << 1));
}

struct ConeGrid {
  double min_x, min_y, bs, h;  // bounding-box corner, box size, h = sqrt(2)/2 bs
  int nce, nrad;               // angular half-count (2 nce + 1 directions), radial intervals
  const double* radial;        // nrad + 1 interval ends, ascending
};

// theta in [0, 2 pi)
IFGF_HD inline double atan2_02pi(double y, double x) {
#pragma clang fp contract(off)
  if (x > 0 && y >= 0) return atan(y / x);
  if (x > 0 && y < 0) return atan(y / x) + 2.0 * M_PI;
  if (x < 0 && y > 0) return atan(y / x) + M_PI;
  if (x < 0 && y == 0) return M_PI;
  if (x < 0 && y < 0) return atan(y / x) + M_PI;
  if (x == 0 && y > 0) return M_PI_2;
  if (x == 0 && y < 0) return 3.0 * M_PI_2;
  return 0.0;
}

IFGF_HD inline void grid_center(const ConeGrid& G, long long m, double& cx, double& cy) {
#pragma clang fp contract(off)
  const long long px = (long long)morton_merge2((uint64_t)m), py = (long long)morton_merge2((uint64_t)m >> 1);
  cx = G.min_x + px * G.bs + 0.5 * G.bs;
  cy = G.min_y + py * G.bs + 0.5 * G.bs;
}

// Non-relevant cone id post * nrad + posr of (x, y) seen from box m, or -1
// outside the cone grid.
IFGF_HD inline long long grid_nonrel(const ConeGrid& G, long long m, double x, double y) {
#pragma clang fp contract(off)
  double cx, cy;
  grid_center(G, m, cx, cy);
  x -= cx;
  y -= cy;
  const double r = sqrt(x * x + y * y);
  if (r == 0.0) return -1;
  const double th = atan2_02pi(y, x);
  // first interval end above r (std::upper_bound over radial[0..nrad])
  int lo = 0, hi = G.nrad + 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (G.radial[mid] > r)
      hi = mid;
    else
      lo = mid + 1;
  }
  const long long posr = (long long)lo - 1;
  const long long post = (long long)(th * G.nce / M_PI);
  if (posr < 0 || posr >= G.nrad || post < 0 || post > 2ll * G.nce) return -1;
  return post * G.nrad + posr;
}

// grid_nonrel, and *near = true when (x, y) lies within a relative 1e-9 of a
// cone boundary (a radial interval end or a direction edge, including the
// grid's outer ends): there a last-bit difference of atan, cos or sin between
// two math libraries (a point computed from them, or its angle) could move it
// to the neighbouring cone. Points on symmetry lines through box centres sit
// exactly on such boundaries. Elsewhere the result cannot differ.
IFGF_HD inline long long grid_nonrel_near(const ConeGrid& G, long long m, double x, double y, bool& near) {
#pragma clang fp contract(off)
  constexpr double kBand = 1e-9;
  double cx, cy;
  grid_center(G, m, cx, cy);
  x -= cx;
  y -= cy;
  const double r = sqrt(x * x + y * y);
  near = r < kBand * G.bs;
  if (r == 0.0) return -1;
  const double th = atan2_02pi(y, x);
  int lo = 0, hi = G.nrad + 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (G.radial[mid] > r)
      hi = mid;
    else
      lo = mid + 1;
  }
  if (lo >= 1 && r - G.radial[lo - 1] < kBand * r) near = true;
  if (lo <= G.nrad && G.radial[lo] - r < kBand * r) near = true;
  const double u = th * G.nce / M_PI;
  const double fu = u - floor(u);
  if (fu < kBand * (1.0 + u) || 1.0 - fu < kBand * (1.0 + u)) near = true;
  const long long posr = (long long)lo - 1;
  const long long post = (long long)u;
  if (posr < 0 || posr >= G.nrad || post < 0 || post > 2ll * G.nce) return -1;
  return post * G.nrad + posr;
}

// Cartesian point of the Chebyshev node (s, t) of cone q of box m.
IFGF_HD inline void grid_cheb2cart(const ConeGrid& G, long long m, long long q, double s, double t, double& x,
                                   double& y) {
#pragma clang fp contract(off)
  const long long posr = q % G.nrad, post = q / G.nrad;
  const double mn = G.radial[posr];
  const double rmin = G.h / G.radial[posr + 1];
  const double dr = G.h / mn - rmin;
  const double r = G.h / ((s + 1.0) / 2.0 * dr + rmin);
  const double da = M_PI / G.nce;
  const double th = (t + 1.0) / 2.0 * da + da * post;
  double cx, cy;
  grid_center(G, m, cx, cy);
  x = r * cos(th) + cx;
  y = r * sin(th) + cy;
}

}  // namespace ifgf
