// Host Householder GMRES (the reference's solver) for operators given as a
// callback; see solver.h.
#include "solver.h"

#include <cmath>
#include <limits>

namespace ifgf {

namespace {

// Reflector H = I - tau [1; v][1; v]^H mapping a (length m) to beta e1, with
// beta = -sign(Re a0) ||a|| (Eigen's makeHouseholder convention). H is unitary
// but not Hermitian for complex a0, so the Krylov vectors and the solution are
// built with H^H = H^-1 (opGMRES.hpp:147-149 and :224-229 apply H as written,
// which for complex data is a consistent but not residual-minimising iteration).
struct Reflector {
  std::vector<cplx> v;
  cplx tau{0.0, 0.0};
};

double make_reflector(const cplx* a, long long m, Reflector& r) {
  double tail = 0.0;
  for (long long i = 1; i < m; i++) tail += std::norm(a[i]);
  const cplx a0 = a[0];
  r.v.assign(m > 0 ? m - 1 : 0, cplx(0.0));
  const double tiny = std::numeric_limits<double>::min();
  if (tail <= tiny && a0.imag() * a0.imag() <= tiny) {
    r.tau = 0.0;
    return a0.real();
  }
  double beta = std::sqrt(std::norm(a0) + tail);
  if (a0.real() >= 0.0) beta = -beta;
  const cplx d = a0 - beta;
  for (long long i = 1; i < m; i++) r.v[i - 1] = a[i] / d;
  r.tau = std::conj((beta - a0) / beta);
  return beta;
}

// a <- H a, or a <- H^H a (= H^-1 a) when adjoint is set.
void reflect(cplx* a, long long m, const Reflector& r, bool adjoint = false) {
  const cplx tau = adjoint ? std::conj(r.tau) : r.tau;
  if (m == 1) {
    a[0] *= (1.0 - tau);
    return;
  }
  if (tau == cplx(0.0)) return;
  cplx s = a[0];
  for (long long i = 1; i < m; i++) s += std::conj(r.v[i - 1]) * a[i];
  const cplx ts = tau * s;
  a[0] -= ts;
  for (long long i = 1; i < m; i++) a[i] -= r.v[i - 1] * ts;
}

// Plane rotation G with G^* (p, q)^T = (r, 0)^T (Eigen's JacobiRotation::makeGivens).
struct Rotation {
  cplx c{1.0, 0.0}, s{0.0, 0.0};
};

Rotation make_rotation(cplx p, cplx q) {
  Rotation g;
  if (q == cplx(0.0)) {
    g.c = p.real() < 0 ? -1.0 : 1.0;
    g.s = 0.0;
    return g;
  }
  if (p == cplx(0.0)) {
    g.c = 0.0;
    g.s = -q / std::abs(q);
    return g;
  }
  const double p1 = std::abs(p.real()) + std::abs(p.imag());
  const double q1 = std::abs(q.real()) + std::abs(q.imag());
  if (p1 >= q1) {
    const cplx ps = p / p1, qs = q / p1;
    const double p2 = std::norm(ps), q2 = std::norm(qs);
    double u = std::sqrt(1.0 + q2 / p2);
    if (p.real() < 0) u = -u;
    g.c = 1.0 / u;
    g.s = -qs * std::conj(ps) * (g.c.real() / p2);
  } else {
    const cplx ps0 = p / q1, qs = q / q1;
    const double p2 = std::norm(ps0), q2 = std::norm(qs);
    double u = q1 * std::sqrt(p2 + q2);
    if (p.real() < 0) u = -u;
    const double pa = std::abs(p);
    const cplx ps = p / pa;
    g.c = pa / u;
    g.s = -std::conj(ps) * (q / u);
  }
  return g;
}

// (a, b) <- G^* (a, b)
void rotate_adj(cplx& a, cplx& b, const Rotation& g) {
  const cplx c = std::conj(g.c), s = -g.s;
  const cplx x = a, y = b;
  a = c * x + std::conj(s) * y;
  b = -s * x + std::conj(c) * y;
}

}  // namespace

bool householder_gmres(const ApplyFn& A, long long n, const cplx* rhs, cplx* x, int restart,
                       double tol, int maxit, int& iters, double& err, std::vector<double>* hist) {
  iters = 0;
  double bn = 0.0;
  for (long long i = 0; i < n; i++) bn += std::norm(rhs[i]);
  if (std::sqrt(bn) <= std::numeric_limits<double>::min()) {
    for (long long i = 0; i < n; i++) x[i] = 0.0;
    err = 0.0;
    return true;
  }
  std::vector<cplx> t(n), v(n), r0(n), dx(n);
  A(x, t.data());
  for (long long i = 0; i < n; i++) r0[i] = rhs[i] - t[i];
  double r0n = 0.0;
  for (auto& q : r0) r0n += std::norm(q);
  r0n = std::sqrt(r0n);
  if (r0n == 0.0) {
    err = 0.0;
    return true;
  }
  const int m = restart;
  std::vector<Reflector> H(m + 1);
  std::vector<cplx> R((size_t)(m + 1) * (m + 1), cplx(0.0));  // R[i*(m+1)+j]
  std::vector<cplx> w(m + 1, cplx(0.0));
  std::vector<Rotation> G(m);
  w[0] = make_reflector(r0.data(), n, H[0]);
  for (int k = 1; k <= m; ++k) {
    ++iters;
    // v = P_1 ... P_k e_k
    std::fill(v.begin(), v.end(), cplx(0.0));
    v[k - 1] = 1.0;
    for (int i = k - 1; i >= 0; --i) reflect(&v[i], n - i, H[i], true);
    A(v.data(), t.data());
    v.swap(t);
    for (int i = 0; i < k; ++i) reflect(&v[i], n - i, H[i]);
    double tn = 0.0;
    for (long long i = k; i < n; i++) tn += std::norm(v[i]);
    if (tn != 0.0 && k < n) {
      make_reflector(&v[k], n - k, H[k]);
      reflect(&v[k], n - k, H[k]);
    }
    for (int i = 0; i < k - 1; ++i) rotate_adj(v[i], v[i + 1], G[i]);
    if (k < n && v[k] != cplx(0.0)) {
      G[k - 1] = make_rotation(v[k - 1], v[k]);
      rotate_adj(v[k - 1], v[k], G[k - 1]);
      rotate_adj(w[k - 1], w[k], G[k - 1]);
    }
    for (int i = 0; i < k; i++) R[(size_t)i * (m + 1) + (k - 1)] = v[i];
    err = std::abs(w[k]) / r0n;
    if (hist) hist->push_back(err);
    const bool stop = (k == n || err < tol || iters >= maxit);
    if (stop || k == m) {
      std::vector<cplx> y(w.begin(), w.begin() + k);
      for (int i = k - 1; i >= 0; --i) {
        for (int j = i + 1; j < k; j++) y[i] -= R[(size_t)i * (m + 1) + j] * y[j];
        y[i] /= R[(size_t)i * (m + 1) + i];
      }
      std::fill(dx.begin(), dx.end(), cplx(0.0));
      for (int i = k - 1; i >= 0; --i) {
        dx[i] += y[i];
        reflect(&dx[i], n - i, H[i], true);
      }
      for (long long i = 0; i < n; i++) x[i] += dx[i];
      if (stop) return k == n || err < tol;
      // restart (opGMRES.hpp:244-261)
      k = 0;
      A(x, t.data());
      for (long long i = 0; i < n; i++) r0[i] = rhs[i] - t[i];
      for (auto& h : H) {
        h.v.clear();
        h.tau = 0.0;
      }
      std::fill(R.begin(), R.end(), cplx(0.0));
      std::fill(w.begin(), w.end(), cplx(0.0));
      w[0] = make_reflector(r0.data(), n, H[0]);
    }
  }
  return false;
}

}  // namespace ifgf
