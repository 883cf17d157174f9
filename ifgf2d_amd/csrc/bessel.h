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

// W simultaneous Clenshaw sums out[q] = sum_{k<N} c[k][q] T_k(t). The W
// recurrences are independent (instruction-level parallelism), and the loop
// is kept rolled so the coefficients are read from the constant table each
// step instead of occupying registers.
template <int N, int W>
IFGF_HD inline void cheb_sums(const double (&c)[N][W], double t, double (&out)[W]) {
  double b1[W], b2[W];
#pragma unroll
  for (int q = 0; q < W; q++) b1[q] = b2[q] = 0.0;
  const double t2 = 2.0 * t;
#pragma unroll 1
  for (int k = N - 1; k >= 1; --k) {
#pragma unroll
    for (int q = 0; q < W; q++) {
      const double b0 = fma(t2, b1[q], c[k][q] - b2[q]);
      b2[q] = b1[q];
      b1[q] = b0;
    }
  }
#pragma unroll
  for (int q = 0; q < W; q++) out[q] = fma(t, b1[q], c[0][q] - b2[q]);
}

// x > 8 branch of hankel0, out of line for the same reason as hankel01_large.
__attribute__((noinline)) IFGF_HD inline zd hankel0_large(double x);

// H0^(1)(x) only (single-layer kernel). x > 0.
IFGF_HD inline zd hankel0(double x) {
  double v[2];
  if (x <= kBesselSplit) {
    cheb_sums(kSmall2, fma(x * x, 1.0 / 32.0, -1.0), v);  // J0, R0
    return zd{v[0], fma(kTwoOverPi * log(0.5 * x), v[0], v[1])};
  }
  return hankel0_large(x);
}

__attribute__((noinline)) IFGF_HD inline zd hankel0_large(double x) {
  double v[2];
  const double ix = 1.0 / x;
  cheb_sums(kLarge2, fma(128.0 * ix, ix, -1.0), v);  // P0, Qt0 at t = 2 (8/x)^2 - 1
  double s, c;
  sincos(x, &s, &c);
  // exp(i(x - pi/4)) = ((c + s) + i(s - c)) / sqrt(2)
  const double amp = sqrt(kTwoOverPi * ix) * 0.70710678118654752440;
  return cmul(zd{v[0], 8.0 * ix * v[1]}, zd{(c + s) * amp, (s - c) * amp});
}

// x > 8 branch of hankel01, kept out of line: the kernels' hot arguments lie
// below the split, and inlining the sincos path raises their register use.
__attribute__((noinline)) IFGF_HD inline void hankel01_large(double x, double ix, zd& h0, zd& h1);

// H0^(1)(x) and H1^(1)(x) together (double-layer / combined-field kernels).
// x > 0, ix = 1/x (callers usually have it from a reciprocal square root).
IFGF_HD inline void hankel01(double x, double ix, zd& h0, zd& h1) {
  double v[4];
  if (x <= kBesselSplit) {
    cheb_sums(kSmall4, fma(x * x, 1.0 / 32.0, -1.0), v);  // J0, J1/x, R0, R1/x
    const double l = kTwoOverPi * log(0.5 * x);
    const double j1 = x * v[1];
    h0 = zd{v[0], fma(l, v[0], v[2])};
    h1 = zd{j1, fma(l, j1, x * v[3]) - kTwoOverPi * ix};
    return;
  }
  hankel01_large(x, ix, h0, h1);
}

IFGF_HD inline void hankel01(double x, zd& h0, zd& h1) { hankel01(x, 1.0 / x, h0, h1); }

__attribute__((noinline)) IFGF_HD inline void hankel01_large(double x, double ix, zd& h0, zd& h1) {
  double v[4];
  cheb_sums(kLarge4, fma(128.0 * ix, ix, -1.0), v);  // P0, Qt0, P1, Qt1
  double s, c;
  sincos(x, &s, &c);
  const double amp = sqrt(kTwoOverPi * ix) * 0.70710678118654752440;
  const zd e0{(c + s) * amp, (s - c) * amp};   // amp * exp(i(x - pi/4))
  const zd e1{e0.y, -e0.x};                    // amp * exp(i(x - 3pi/4))
  h0 = cmul(zd{v[0], 8.0 * ix * v[1]}, e0);
  h1 = cmul(zd{v[2], 8.0 * ix * v[3]}, e1);
}

}  // namespace ifgf
