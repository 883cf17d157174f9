// Geometry of the closed curve and one column of the rectangular-polar (RP)
// singular / near-singular precomputation, shared by the host setup and the
// gfx950 kernel that builds the correction on the device.
#pragma once

#include <cmath>

#include "green.h"

namespace ifgf {

constexpr int kMaxStarModes = 64;

// Closed parametrised curve, t in [0, 2 pi) (Geometry/Curve.hpp:21-146).
//   type 0 circle (cx, cy, R); 1 kite x = A cos t + B cos 2t + C, y = D sin t;
//   2 star (cx, cy, R, nmodes, a[0..nmodes), b[0..nmodes)),
//     r(t) = R (1 + sum_m a_m cos mt + b_m sin mt).
struct Curve {
  int type = 0;
  int nm = 0;
  double cx = 0, cy = 0, R = 1;
  double A = 1, B = 0.65, C = -0.65, D = 1.5;
  double am[kMaxStarModes] = {}, bm[kMaxStarModes] = {};

  IFGF_HD double x(double t) const {
    if (type == 0) return R * cos(t) + cx;
    if (type == 1) return A * cos(t) + B * cos(2.0 * t) + C;
    double r = 1.0;
    for (int m = 0; m < nm; m++) r += am[m] * cos(m * t) + bm[m] * sin(m * t);
    return cx + R * r * cos(t);
  }
  IFGF_HD double y(double t) const {
    if (type == 0) return R * sin(t) + cy;
    if (type == 1) return D * sin(t);
    double r = 1.0;
    for (int m = 0; m < nm; m++) r += am[m] * cos(m * t) + bm[m] * sin(m * t);
    return cy + R * r * sin(t);
  }
  IFGF_HD void star_rrp(double t, double& r, double& rp) const {
    double a = 1.0, d = 0.0;
    for (int m = 0; m < nm; m++) {
      a += am[m] * cos(m * t) + bm[m] * sin(m * t);
      d += -am[m] * m * sin(m * t) + bm[m] * m * cos(m * t);
    }
    r = R * a;
    rp = R * d;
  }
  IFGF_HD double xp(double t) const {
    if (type == 0) return -R * sin(t);
    if (type == 1) return -A * sin(t) - 2.0 * B * sin(2.0 * t);
    double r, rp;
    star_rrp(t, r, rp);
    return rp * cos(t) - r * sin(t);
  }
  IFGF_HD double yp(double t) const {
    if (type == 0) return R * cos(t);
    if (type == 1) return D * cos(t);
    double r, rp;
    star_rrp(t, r, rp);
    return rp * sin(t) + r * cos(t);
  }
  // x(a) - x(a - d), y(a) - y(a - d) from trigonometric difference identities,
  // accurate to rounding relative to the (small) difference itself.
  IFGF_HD void diff(double a, double d, double& dx, double& dy) const {
    const double mid = a - 0.5 * d;
    const double s1 = sin(0.5 * d);
    if (type == 0) {
      dx = -2.0 * R * sin(mid) * s1;
      dy = 2.0 * R * cos(mid) * s1;
      return;
    }
    if (type == 1) {
      dx = -2.0 * A * sin(mid) * s1 - 2.0 * B * sin(2.0 * mid) * sin(d);
      dy = 2.0 * D * cos(mid) * s1;
      return;
    }
    // star: x = cx + r(t) cos t, r(a) - r(a - d) by the same identities
    double ra = 1.0, drr = 0.0;
    for (int m = 0; m < nm; m++) {
      ra += am[m] * cos(m * a) + bm[m] * sin(m * a);
      const double sm = sin(0.5 * m * d);
      drr += -2.0 * am[m] * sin(m * mid) * sm + 2.0 * bm[m] * cos(m * mid) * sm;
    }
    ra *= R;
    drr *= R;
    const double b = a - d;
    dx = ra * (-2.0 * sin(mid) * s1) + drr * cos(b);
    dy = ra * (2.0 * cos(mid) * s1) + drr * sin(b);
  }
};

// RP change of variables w(t) and w'(t), utils/Quadrature.hpp:73-102.
IFGF_HD inline double rp_vfun(double t, double p) {
  const double u = t / M_PI - 1.0;
  return (0.5 - 1.0 / p) * u * u * u + u / p + 0.5;
}
IFGF_HD inline double rp_vpfun(double t, double p) {
  const double u = t / M_PI - 1.0;
  return (3.0 / M_PI) * (0.5 - 1.0 / p) * u * u + 1.0 / (p * M_PI);
}
IFGF_HD inline double rp_wcov(double t, double p) {
  const double a = pow(rp_vfun(t, p), p), b = pow(rp_vfun(2 * M_PI - t, p), p);
  return 2.0 * M_PI * a / (a + b);
}
IFGF_HD inline double rp_wpcov(double t, double p) {
  const double v1 = rp_vfun(t, p), v2 = rp_vfun(2 * M_PI - t, p);
  const double a = pow(v1, p), b = pow(v2, p);
  return 2.0 * M_PI * p * rp_vpfun(t, p) * pow(v1, p - 1.0) * pow(v2, p - 1.0) / ((a + b) * (a + b));
}
IFGF_HD inline int rp_sgn(double v) { return (0.0 < v) - (v < 0.0); }

// xi_cov_{center,left,right} and derivatives, utils/Quadrature.hpp:118-147.
// off = xi - xi_s, the offset from the singular location xi_s (alpha, -1 or
// +1), returned without cancellation.
IFGF_HD inline void rp_map(int loc, double t, double alpha, double p, double& xi, double& dxi, double& off) {
  if (loc == 0) {
    off = ((rp_sgn(t) - alpha) / M_PI) * rp_wcov(M_PI * fabs(t), p);
    xi = alpha + off;
    dxi = (1.0 - alpha * rp_sgn(t)) * rp_wpcov(M_PI * fabs(t), p);
  } else if (loc == -1) {
    const double u = (t + 1.0) / 2.0;
    off = (2.0 / M_PI) * rp_wcov(M_PI * fabs(u), p);
    xi = -1.0 + off;
    dxi = rp_wpcov(M_PI * fabs(u), p) * rp_sgn(u);
  } else {
    const double u = (t - 1.0) / 2.0;
    off = -((2.0 / M_PI) * rp_wcov(M_PI * fabs(u), p));
    xi = 1.0 + off;
    dxi = -rp_wpcov(M_PI * fabs(u), p) * rp_sgn(u);
  }
}

// One column of the precomputation matrix: Np Chebyshev moments of the kernel
// against the patch [t1, t2], integrated with the nq-node RP rule (fn, fw)
// around tsing (single_patch_point_compute_precomputations,
// ForwardMap.tpp:287-389). The target is the curve point at parameter ttgt;
// the separation x - y is formed from parameter differences (Curve::diff), so
// the double-layer numerator (x - y).n_y keeps full relative accuracy as
// y -> x.
IFGF_HD inline void rp_column(const Curve& c, int F, double k, double eta, double p, double t1, double t2,
                              double tsing, double ttgt, int Np, const double* fn, const double* fw, int nq,
                              zd* out) {
  const double alpha = ((1.0 - -1.0) / (t2 - t1)) * (tsing - t1) + -1.0;
  int loc = 0;
  if (fabs(-1.0 - alpha) < 1e-5) loc = -1;
  else if (fabs(1.0 - alpha) < 1e-5) loc = 1;
  for (int n = 0; n < Np; n++) out[n] = zd{0.0, 0.0};
  // parameter of the singular location the change of variables is centred on
  const double ts = loc == 0 ? tsing : (loc < 0 ? t1 : t2);
  double tt = ttgt - ts;  // target offset from it (wrapped to (-pi, pi])
  if (tt > M_PI) tt -= 2.0 * M_PI;
  if (tt < -M_PI) tt += 2.0 * M_PI;
  for (int i = 0; i < nq; i++) {
    double xi, dxi, off;
    rp_map(loc, fn[i], alpha, p, xi, dxi, off);
    const double g = ((t2 - t1) / 2.0) * (xi + 1.0) + t1;
    const double n0 = c.yp(g), n1 = -c.xp(g);
    const double cj = sqrt(n0 * n0 + n1 * n1);
    double dx, dy;
    c.diff(ttgt, tt - ((t2 - t1) / 2.0) * off, dx, dy);
    const zd gr = green_d_rt(F, dx, dy, n0 / cj, n1 / cj, k, eta);
    const double s = ((t2 - t1) / 2.0) * cj * dxi * fw[i];
    const zd v{gr.x * s, gr.y * s};
    // Chebyshev moments T_n(xi), n < Np (ForwardMap.tpp:373-385)
    double tm = 1.0, tc = xi;
    out[0] = cadd(out[0], v);
    out[1] = cadd(out[1], cscale(v, tc));
    for (int n = 2; n < Np; n++) {
      const double tn = 2.0 * xi * tc - tm;
      out[n] = cadd(out[n], cscale(v, tn));
      tm = tc;
      tc = tn;
    }
  }
}

// Near-singular test of the point (xq, yq) against the patch [a, b]
// (determine_patch_near_singular_points, ForwardMap.tpp:224-285): true when
// the point lies within delta of the patch, with ts the closest parameter --
// a if |x(a) - x_q| < delta, else b if that holds, else the interior
// minimiser of the distance by golden-section search to 1e-12 (the
// reference's Boost Brent, BrentsMethod.hpp:42-156, finds the same minimum).
// Host setup and the device search share this code.
IFGF_HD inline bool near_singular_param(const Curve& c, double a, double b, double xq, double yq, double delta,
                                        double& ts) {
  auto dist = [&](double t) {
    const double dx = c.x(t) - xq, dy = c.y(t) - yq;
    return sqrt(fabs(dx * dx + dy * dy));
  };
  if (dist(a) < delta) {
    ts = a;
    return true;
  }
  if (dist(b) < delta) {
    ts = b;
    return true;
  }
  const double gr = 0.5 * (sqrt(5.0) - 1.0);
  double lo = a, hi = b, c1 = hi - gr * (hi - lo), c2 = lo + gr * (hi - lo);
  double f1 = dist(c1), f2 = dist(c2);
  for (int it = 0; it < 200 && (hi - lo) > 1e-12 * (1.0 + fabs(lo)); it++) {
    if (f1 < f2) {
      hi = c2;
      c2 = c1;
      f2 = f1;
      c1 = hi - gr * (hi - lo);
      f1 = dist(c1);
    } else {
      lo = c1;
      c1 = c2;
      f1 = f2;
      c2 = lo + gr * (hi - lo);
      f2 = dist(c2);
    }
  }
  ts = 0.5 * (lo + hi);
  return dist(ts) < delta;
}

}  // namespace ifgf
