// FP64 Hankel functions H0^(1)(x), H1^(1)(x) for real x > 0, written for both
// the host setup code and the gfx950 kernels.
//
// Replaces sp_bessel::hankelH1(0|1, k r)
// (complex_bessel-master/include/complex_bessel_bits/besselFunctions.h:312-341,
// called from ForwardMap/GreenFunctions.hpp:49,56). The coefficient tables come
// from gen_bessel_coeffs.py (mpmath); the two regions and the fitted functions
// are described there. Only published mathematics is used: the ascending series
// structure of A&S 9.1.10-9.1.11 for x <= 8 and the Hankel amplitude/phase form
// of A&S 9.2.5-9.2.6 for x > 8.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "bessel_coeffs.h"
#include "zmath.h"

namespace ifgf {

constexpr double kTwoOverPi = 0.63661977236758134308;   // 2/pi
constexpr double kBesselSplit = 8.0;

// Clenshaw summation of sum_{k<N} c_k T_k(t).
template <int N>
IFGF_HD inline double cheb_sum(const double (&c)[N], double t) {
  double b1 = 0.0, b2 = 0.0;
  const double t2 = 2.0 * t;
#pragma unroll
  for (int k = N - 1; k >= 1; --k) {
    const double b0 = fma(t2, b1, c[k] - b2);
    b2 = b1;
    b1 = b0;
  }
  return fma(t, b1, c[0] - b2);
}

// H0^(1)(x) only (single-layer kernel). x > 0.
IFGF_HD inline zd hankel0(double x) {
  if (x <= kBesselSplit) {
    const double t = fma(x * x, 1.0 / 32.0, -1.0);
    const double j0 = cheb_sum(kJ0Small, t);
    const double r0 = cheb_sum(kR0Small, t);
    const double y0 = fma(kTwoOverPi * log(0.5 * x), j0, r0);
    return zd{j0, y0};
  }
  const double ix = 1.0 / x;
  const double z = 64.0 * ix * ix;
  const double t = fma(2.0, z, -1.0);
  const double p0 = cheb_sum(kP0Large, t);
  const double q0 = 8.0 * ix * cheb_sum(kQ0Large, t);
  double s, c;
  sincos(x, &s, &c);
  // exp(i(x - pi/4)) = ((c + s) + i(s - c)) / sqrt(2)
  const double amp = sqrt(kTwoOverPi * ix) * 0.70710678118654752440;
  const zd ph{(c + s) * amp, (s - c) * amp};
  return cmul(zd{p0, q0}, ph);
}

// H0^(1)(x) and H1^(1)(x) together (double-layer / combined-field kernels). x > 0.
IFGF_HD inline void hankel01(double x, zd& h0, zd& h1) {
  if (x <= kBesselSplit) {
    const double t = fma(x * x, 1.0 / 32.0, -1.0);
    const double j0 = cheb_sum(kJ0Small, t);
    const double r0 = cheb_sum(kR0Small, t);
    const double j1 = x * cheb_sum(kJ1Small, t);
    const double r1 = x * cheb_sum(kR1Small, t);
    const double l = kTwoOverPi * log(0.5 * x);
    h0 = zd{j0, fma(l, j0, r0)};
    h1 = zd{j1, fma(l, j1, r1) - kTwoOverPi / x};
    return;
  }
  const double ix = 1.0 / x;
  const double z = 64.0 * ix * ix;
  const double t = fma(2.0, z, -1.0);
  const double p0 = cheb_sum(kP0Large, t);
  const double q0 = 8.0 * ix * cheb_sum(kQ0Large, t);
  const double p1 = cheb_sum(kP1Large, t);
  const double q1 = 8.0 * ix * cheb_sum(kQ1Large, t);
  double s, c;
  sincos(x, &s, &c);
  const double amp = sqrt(kTwoOverPi * ix) * 0.70710678118654752440;
  const zd e0{(c + s) * amp, (s - c) * amp};   // amp * exp(i(x - pi/4))
  const zd e1{e0.y, -e0.x};                    // amp * exp(i(x - 3pi/4))
  h0 = cmul(zd{p0, q0}, e0);
  h1 = cmul(zd{p1, q1}, e1);
}

}  // namespace ifgf
