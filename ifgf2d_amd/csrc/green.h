// Layer-potential kernels of the 2D Helmholtz boundary-integral operator, for
// host setup code and gfx950 kernels alike.
//
// Counterpart of GF<Formulation> / GF_far (ForwardMap/GreenFunctions.hpp:35-109)
// and factorization() (GreenFunctions.hpp:125-132). Convention: x = target,
// y = source, n = unit normal at the source y.
//   single layer   G(x,y)      = (i/4) H0(k r)
//   double layer   dG/dn_y     = (ik/4) H1(k r) (x-y).n_y / r
//   combined field dG/dn_y - i eta G
// The single layer is the reference's (symmetric in x,y). The double-layer and
// combined forms follow the mathematics; the reference's versions are commented
// out or miss the 1/r factor (GreenFunctions.hpp:55-57,68-75).
#pragma once

#include "bessel.h"

namespace ifgf {

enum Formulation : int { kSingleLayer = 0, kDoubleLayer = 1, kCombined = 2 };

// Kernel from the separation d = x - y and the unit source normal n. One
// reciprocal square root gives r, 1/r and 1/(k r).
template <int F>
IFGF_HD inline zd green_d(double dx, double dy, double nx, double ny, double k, double eta) {
  const double r2 = dx * dx + dy * dy;
  if (r2 < 1e-28) return zd{0.0, 0.0};  // r < 1e-14, GreenFunctions.hpp:97-100
  const double ri = rsqrt_hd(r2);
  const double r = r2 * ri;
  if (F == kSingleLayer) {
    const zd h0 = hankel0(k * r);
    return zd{-0.25 * h0.y, 0.25 * h0.x};
  } else {
    zd h0, h1;
    hankel01(k * r, ri / k, h0, h1);
    const double c = 0.25 * k * (dx * nx + dy * ny) * ri;
    zd out{-c * h1.y, c * h1.x};  // i * c * H1
    if (F == kCombined) {
      out.x = fma(0.25 * eta, h0.x, out.x);  // -i eta (i/4) H0 = (eta/4) H0
      out.y = fma(0.25 * eta, h0.y, out.y);
    }
    return out;
  }
}

template <int F>
IFGF_HD inline zd green(double tx, double ty, double sx, double sy, double nx, double ny, double k,
                        double eta) {
  return green_d<F>(tx - sx, ty - sy, nx, ny, k, eta);
}

IFGF_HD inline zd green_d_rt(int F, double dx, double dy, double nx, double ny, double k, double eta) {
  if (F == kSingleLayer) return green_d<kSingleLayer>(dx, dy, nx, ny, k, eta);
  if (F == kDoubleLayer) return green_d<kDoubleLayer>(dx, dy, nx, ny, k, eta);
  return green_d<kCombined>(dx, dy, nx, ny, k, eta);
}

IFGF_HD inline zd green_rt(int F, double tx, double ty, double sx, double sy, double nx, double ny,
                           double k, double eta) {
  if (F == kSingleLayer) return green<kSingleLayer>(tx, ty, sx, sy, nx, ny, k, eta);
  if (F == kDoubleLayer) return green<kDoubleLayer>(tx, ty, sx, sy, nx, ny, k, eta);
  return green<kCombined>(tx, ty, sx, sy, nx, ny, k, eta);
}

// exp(i k r) / sqrt(r), GreenFunctions.hpp:125-132.
IFGF_HD inline zd centre_factor(double r, double k) {
  double s, c;
  sincos(k * r, &s, &c);
  const double a = 1.0 / sqrt(r);
  return zd{c * a, s * a};
}

// 1 / (exp(i k r) / sqrt(r)) = exp(-i k r) sqrt(r)
IFGF_HD inline zd inv_centre_factor(double r, double k) {
  double s, c;
  sincos(k * r, &s, &c);
  const double a = sqrt(r);
  return zd{c * a, -s * a};
}

}  // namespace ifgf
