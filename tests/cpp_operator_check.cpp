// The C ABI driven from C++ on the GPU, the way a maintainer's binding in
// the reference would use it (ForwardMap::solve, ForwardMap.tpp:658-697):
// the forward map wrapped as a LinearOperator-style object (y = A x on
// std::complex<double> vectors, the mapping of LinearOperator.hpp:54), solved
// (a) by the library's restatement of the reference's host GMRES
// (opGMRES.hpp:68-269) over that operator and (b) by the device GMRES, for
// the sound-soft unit circle at k = 10 (40 patches x 10 nodes, CFIE, 3
// levels) and the plane wave d = (1, 0); the scattered field at 64 points on
// r = 3 (propagate_unacc, ForwardMap.tpp:702-727) is compared with the Mie
// series
//   u(r, th) = -sum_m i^m J_m(k) / H_m(k) H_m(k r) e^{i m th},
// computed here with the C++17 Bessel functions. 9 x 9 cones: field within
// 1e-5; 5 x 5 cones: within 5e-4 (the cone interpolation error). Built by
// __graft_entry__.build() (g++, no HIP headers); run by
// tests/test_gpu_cpp_binding.py. Prints one line per case and "ok", exit 0
// on success.
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <vector>

#include "ifgf2d.h"

using cd = std::complex<double>;
using Vec = std::vector<cd>;

namespace {

const double kPi = 3.14159265358979323846;

// LinearOperator (LinearOperator.hpp:27-60) without Eigen: rows, cols and a
// mapping x -> y.
struct LinearOperator {
  long long n;
  std::function<void(const Vec&, Vec&)> map;
  Vec operator*(const Vec& x) const {
    Vec y(x.size());
    map(x, y);
    return y;
  }
};

ifgf_plan* make_plan(int Ps, int Pang) {
  static double lims[41];
  for (int p = 0; p <= 40; p++) lims[p] = 2.0 * kPi * p / 40;
  static const double circle[3] = {0.0, 0.0, 1.0};
  const double delta = 0.75 * 2.0 * std::sin(kPi / 40);
  ifgf_surface s{IFGF_CURVE_CIRCLE, circle, 3, lims, 40, /*Np*/ 10, delta, /*p*/ 6.0};
  ifgf_params prm{IFGF_COMBINED, IFGF_FP64, 10.0, /*eta*/ 10.0, Ps, Pang, /*nlevels*/ 3, /*nrhs_max*/ 1, 0};
  ifgf_plan* plan = nullptr;
  if (ifgf_plan_create(&s, &prm, &plan) != IFGF_OK) throw std::runtime_error(ifgf_last_error());
  return plan;
}

// H_m(x) = J_m(x) + i Y_m(x) for integer m of either sign
cd hankel(int m, double x) {
  const int a = m < 0 ? -m : m;
  const cd h(std::cyl_bessel_j((double)a, x), std::cyl_neumann((double)a, x));
  return (m < 0 && (a % 2)) ? -h : h;
}

cd mie(double r, double th, double k) {
  cd u = 0.0;
  const int M = (int)k + 40;
  for (int m = -M; m <= M; m++) {
    const cd im = std::pow(cd(0.0, 1.0), m);
    const double jm = (m < 0 && ((-m) % 2)) ? -std::cyl_bessel_j((double)(-m), k) : std::cyl_bessel_j((double)(m < 0 ? -m : m), k);
    u -= im * jm / hankel(m, k) * hankel(m, k * r) * std::exp(cd(0.0, m * th));
  }
  return u;
}

double rel(const Vec& a, const Vec& b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); i++) {
    num += std::norm(a[i] - b[i]);
    den += std::norm(b[i]);
  }
  return std::sqrt(num / den);
}

void host_apply(void* ctx, const double* x, double* y) {
  const LinearOperator& A = *static_cast<const LinearOperator*>(ctx);
  Vec xv(reinterpret_cast<const cd*>(x), reinterpret_cast<const cd*>(x) + A.n);
  const Vec yv = A * xv;
  for (long long i = 0; i < A.n; i++) reinterpret_cast<cd*>(y)[i] = yv[i];
}

}  // namespace

int main() {
  const double k = 10.0;
  Vec ref(64);
  std::vector<double> tx(64), ty(64);
  for (int t = 0; t < 64; t++) {
    const double th = 2.0 * kPi * t / 64;
    tx[t] = 3.0 * std::cos(th);
    ty[t] = 3.0 * std::sin(th);
    ref[t] = mie(3.0, th, k);
  }
  const struct {
    int Ps, Pang;
    double bar;
  } cases[] = {{9, 9, 1e-5}, {5, 5, 5e-4}};
  for (const auto& c : cases) {
    ifgf_plan* plan = make_plan(c.Ps, c.Pang);
    const long long n = ifgf_plan_size(plan);
    std::vector<double> x(n), y(n);
    if (ifgf_plan_points(plan, x.data(), y.data(), nullptr, nullptr, nullptr) != IFGF_OK) return 1;
    Vec rhs(n);
    for (long long i = 0; i < n; i++) rhs[i] = -std::exp(cd(0.0, k * x[i]));
    // the mapping of LinearOperator: the GPU forward map on host vectors
    LinearOperator A{n, [plan](const Vec& xin, Vec& out) {
                       if (ifgf_apply(plan, reinterpret_cast<const double*>(xin.data()),
                                      reinterpret_cast<double*>(out.data()), 1, (long long)xin.size(), 0,
                                      nullptr) != IFGF_OK)
                         throw std::runtime_error(ifgf_plan_error(plan));
                     }};
    // (a) host GMRES over the operator, (b) device GMRES
    Vec dh(n, 0.0), dd(n, 0.0);
    int ith = 0, itd = 0;
    double errh = 0.0, relres = 0.0;
    const int rch = ifgf_gmres_host(host_apply, &A, n, reinterpret_cast<const double*>(rhs.data()),
                                    reinterpret_cast<double*>(dh.data()), 50, 1e-8, 500, &ith, &errh);
    const int rcd = ifgf_gmres(plan, reinterpret_cast<const double*>(rhs.data()), reinterpret_cast<double*>(dd.data()),
                               50, 1e-8, 500, &itd, &relres, nullptr, 0);
    if (rch != IFGF_OK || rcd != IFGF_OK) {
      std::printf("GMRES failed: host %d device %d (%s)\n", rch, rcd, ifgf_last_error());
      return 1;
    }
    // true residual of the host solution through the operator object
    const double rh = rel(A * dh, rhs);
    Vec uh(64), ud(64);
    if (ifgf_propagate(plan, tx.data(), ty.data(), 64, reinterpret_cast<const double*>(dh.data()),
                       reinterpret_cast<double*>(uh.data())) != IFGF_OK ||
        ifgf_propagate(plan, tx.data(), ty.data(), 64, reinterpret_cast<const double*>(dd.data()),
                       reinterpret_cast<double*>(ud.data())) != IFGF_OK)
      return 1;
    const double eh = rel(uh, ref), ed = rel(ud, ref), ddiff = rel(dh, dd);
    std::printf("circle k=10 %dx%d: host GMRES %d it (relres %.2e), device GMRES %d it (relres %.2e), "
                "densities %.2e apart, Mie error host %.3e device %.3e (bar %.0e)\n",
                c.Ps, c.Pang, ith, rh, itd, relres, ddiff, eh, ed, c.bar);
    ifgf_plan_destroy(plan);
    if (!(rh < 1e-7 && relres < 1e-7 && ddiff < 1e-6 && eh < c.bar && ed < c.bar)) return 1;
  }
  std::printf("ok\n");
  return 0;
}
