// Minimal complex arithmetic shared by host setup code and gfx950 kernels.
// Complex numbers are stored interleaved (re, im), the layout of
// std::complex<double> / Eigen::VectorXcd that the reference passes around.
#pragma once

#include <hip/hip_runtime.h>

#define IFGF_HD __host__ __device__

namespace ifgf {

struct zd {
  double x, y;
};

struct zf {
  float x, y;
};

// 1/sqrt(x): the device's reciprocal square root; on the host 1/sqrt.
IFGF_HD inline double rsqrt_hd(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return rsqrt(x);
#else
  return 1.0 / sqrt(x);
#endif
}

IFGF_HD inline zd cadd(zd a, zd b) { return zd{a.x + b.x, a.y + b.y}; }
IFGF_HD inline zd csub(zd a, zd b) { return zd{a.x - b.x, a.y - b.y}; }
IFGF_HD inline zd cmul(zd a, zd b) { return zd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
IFGF_HD inline zd cscale(zd a, double s) { return zd{a.x * s, a.y * s}; }
// a += b * c
IFGF_HD inline void cfma(zd& a, zd b, zd c) {
  a.x = fma(b.x, c.x, fma(-b.y, c.y, a.x));
  a.y = fma(b.x, c.y, fma(b.y, c.x, a.y));
}
// a / b
IFGF_HD inline zd cdiv(zd a, zd b) {
  const double d = 1.0 / (b.x * b.x + b.y * b.y);
  return zd{(a.x * b.x + a.y * b.y) * d, (a.y * b.x - a.x * b.y) * d};
}

// The same operations for either precision (Z = zd or zf), used by the kernels
// that run in FP64 on a plain plan and in FP32 on a mixed-precision plan.
template <typename R>
struct complex_of;
template <>
struct complex_of<double> {
  using type = zd;
};
template <>
struct complex_of<float> {
  using type = zf;
};
template <typename Z>
IFGF_HD inline Z zadd(Z a, Z b) {
  return Z{a.x + b.x, a.y + b.y};
}
template <typename Z>
IFGF_HD inline Z zmul(Z a, Z b) {
  return Z{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename Z>
IFGF_HD inline void zfma(Z& a, Z b, Z c) {
  a.x = fma(b.x, c.x, fma(-b.y, c.y, a.x));
  a.y = fma(b.x, c.y, fma(b.y, c.x, a.y));
}
template <typename Z>
IFGF_HD inline Z zcast(zd a) {
  return Z{(decltype(Z::x))a.x, (decltype(Z::y))a.y};
}
template <typename Z>
IFGF_HD inline zd zwiden(Z a) {
  return zd{(double)a.x, (double)a.y};
}

}  // namespace ifgf
