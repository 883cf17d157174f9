// Restarted GMRES on the GPU for the plan's forward map (ifgf_gmres).
//
// Replaces ForwardMap::solve's Eigen::opGMRES<LinearOperator>
// (ForwardMap.tpp:658-697; Householder Arnoldi in opGMRES.hpp:68-269) with an
// Arnoldi process that keeps the Krylov basis V (n x (m+1) complex, column
// major) in HBM and orthogonalises by classical Gram-Schmidt with one
// re-orthogonalisation pass (CGS2), all in FP64. The projections stay on the
// device; one (j+2)-value copy per iteration brings the new Hessenberg column
// to the host, where the Givens rotations and the residual estimate
// |w(k)| / ||b - A x0|| (opGMRES.hpp:207) are updated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "plan_internal.h"

using namespace ifgf;

namespace {

constexpr int kChunks = 512;

// part[col * kChunks + blockIdx.x] = sum over this block's rows of conj(V[:,col]) * w
__global__ void __launch_bounds__(256)
    k_mdot(long long n, const zd* __restrict__ V, long long ldv, const zd* __restrict__ w,
           zd* __restrict__ part) {
  const int col = blockIdx.y;
  const zd* v = V + (size_t)col * ldv;
  zd acc{0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const zd a = v[i], b = w[i];
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
  }
  __shared__ double sx[256], sy[256];
  sx[threadIdx.x] = acc.x;
  sy[threadIdx.x] = acc.y;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sx[threadIdx.x] += sx[threadIdx.x + s];
      sy[threadIdx.x] += sy[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)col * kChunks + blockIdx.x] = zd{sx[0], sy[0]};
}

// h[col] (+)= sum of the chunk partials; one block per column
__global__ void __launch_bounds__(256) k_reduce(const zd* __restrict__ part, zd* __restrict__ h, int accumulate) {
  const int col = blockIdx.x;
  __shared__ double sx[256], sy[256];
  double ax = 0.0, ay = 0.0;
  for (int c = threadIdx.x; c < kChunks; c += 256) {
    ax += part[(size_t)col * kChunks + c].x;
    ay += part[(size_t)col * kChunks + c].y;
  }
  sx[threadIdx.x] = ax;
  sy[threadIdx.x] = ay;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sx[threadIdx.x] += sx[threadIdx.x + s];
      sy[threadIdx.x] += sy[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (accumulate) h[col] = cadd(h[col], zd{sx[0], sy[0]});
    else h[col] = zd{sx[0], sy[0]};
  }
}

// w -= sign * sum_{j < ncol} V[:,j] h[j]
__global__ void __launch_bounds__(256)
    k_update(long long n, int ncol, const zd* __restrict__ V, long long ldv, const zd* __restrict__ h,
             double sign, zd* __restrict__ w) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  zd acc = w[i];
  for (int j = 0; j < ncol; j++) {
    const zd c = cscale(h[j], -sign);
    cfma(acc, c, V[(size_t)j * ldv + i]);
  }
  w[i] = acc;
}

// dst = src / sqrt(|nrm|) (nrm = ||src||^2 computed on the device)
__global__ void __launch_bounds__(256)
    k_normalize(long long n, const zd* __restrict__ src, const zd* __restrict__ nrm, zd* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = 1.0 / sqrt(nrm[0].x);
  dst[i] = cscale(src[i], s);
}

// r = b - r
__global__ void __launch_bounds__(256) k_residual(long long n, const zd* __restrict__ b, zd* __restrict__ r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = csub(b[i], r[i]);
}

struct Givens {
  std::complex<double> c, s;
};

// complex Givens rotation zeroing b in (a, b): [c s; -conj(s) c] with real c
Givens make_givens(std::complex<double> a, std::complex<double> b) {
  Givens g;
  const double aa = std::abs(a), bb = std::abs(b);
  if (bb == 0.0) {
    g.c = 1.0;
    g.s = 0.0;
  } else if (aa == 0.0) {
    g.c = 0.0;
    g.s = std::conj(b) / bb;
  } else {
    const double r = std::hypot(aa, bb);
    g.c = aa / r;
    g.s = (a / aa) * std::conj(b) / r;
  }
  return g;
}
void apply_givens(const Givens& g, std::complex<double>& a, std::complex<double>& b) {
  const std::complex<double> t = g.c * a + g.s * b;
  b = -std::conj(g.s) * a + g.c * b;
  a = t;
}

// Device workspace of one GMRES cycle of restart m on vectors of length n.
struct Krylov {
  DevBuf<zd> V, w, tmp, part, h;
  Krylov(long long n, int m) {
    V.alloc((size_t)n * (m + 1));
    w.alloc(n);
    tmp.alloc(n);
    part.alloc((size_t)kChunks * (m + 2));
    h.alloc(2 * (m + 2));  // [0, m+2): projections; [m+2, 2m+4): second CGS pass
  }
};

struct Solver {
  ifgf_plan* P;
  hipStream_t st;
  long long n;
  int nb;
  Solver(ifgf_plan* p) : P(p), st(p->stream), n(p->geo.N), nb((int)((p->geo.N + 255) / 256)) {}

  void apply(const zd* x, zd* y, bool fp64) {
    const int rc = apply_device(P, x, y, st, fp64);
    if (rc) throw std::runtime_error(P->err);
  }
  // ||v||^2 into dst[0].x (device)
  void norm2(Krylov& K, const zd* v, zd* dst) {
    hipLaunchKernelGGL(k_mdot, dim3(kChunks, 1), dim3(256), 0, st, n, v, n, v, K.part.p);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, K.part.p, dst, 0);
  }
  double norm(Krylov& K, const zd* v) {
    norm2(K, v, K.h.p);
    std::vector<std::complex<double>> hv;
    fetch(K.h.p, 1, hv);
    return std::sqrt(hv[0].real());
  }
  void fetch(const zd* src, int cnt, std::vector<std::complex<double>>& out) {
    out.resize(cnt);
    HIP_TRY(hipMemcpyAsync(out.data(), src, cnt * sizeof(zd), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  // r = b - A x
  void residual(const zd* b, const zd* x, zd* r, bool fp64) {
    apply(x, r, fp64);
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, st, n, b, r);
  }
  // Appends column j+1 of the Arnoldi basis: w is orthogonalised against
  // V[:, 0..j] by CGS2 and normalised into V[:, j+1]; returns the Hessenberg
  // column h[0..j+1] on the host.
  void arnoldi_step(Krylov& K, int m, int j, std::vector<std::complex<double>>& hv) {
    hipLaunchKernelGGL(k_mdot, dim3(kChunks, j + 1), dim3(256), 0, st, n, K.V.p, n, K.w.p, K.part.p);
    hipLaunchKernelGGL(k_reduce, dim3(j + 1), dim3(256), 0, st, K.part.p, K.h.p, 0);
    hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, n, j + 1, K.V.p, n, K.h.p, 1.0, K.w.p);
    hipLaunchKernelGGL(k_mdot, dim3(kChunks, j + 1), dim3(256), 0, st, n, K.V.p, n, K.w.p, K.part.p);
    hipLaunchKernelGGL(k_reduce, dim3(j + 1), dim3(256), 0, st, K.part.p, K.h.p + m + 2, 0);
    hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, n, j + 1, K.V.p, n, K.h.p + m + 2, 1.0, K.w.p);
    hipLaunchKernelGGL(k_reduce, dim3(j + 1), dim3(256), 0, st, K.part.p, K.h.p, 1);  // h += h2
    norm2(K, K.w.p, K.h.p + j + 1);
    hipLaunchKernelGGL(k_normalize, dim3(nb), dim3(256), 0, st, n, K.w.p, K.h.p + j + 1, K.V.p + (size_t)(j + 1) * n);
    fetch(K.h.p, j + 2, hv);
    hv[j + 1] = std::sqrt(hv[j + 1].real());
  }
  // x += B y for the first kk columns of B (ld n)
  void update(Krylov& K, const zd* B, int kk, const std::vector<std::complex<double>>& y, zd* x) {
    HIP_TRY(hipMemcpyAsync(K.h.p, y.data(), kk * sizeof(zd), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, n, kk, B, n, K.h.p, -1.0, x);
  }
};

// Upper-triangular solve R y = g (R = the rotated Hessenberg, row stride m).
std::vector<std::complex<double>> back_substitute(const std::vector<std::complex<double>>& H, int m,
                                                  const std::vector<std::complex<double>>& g, int kk) {
  std::vector<std::complex<double>> y(kk);
  for (int i = kk - 1; i >= 0; i--) {
    std::complex<double> s = g[i];
    for (int q = i + 1; q < kk; q++) s -= H[(size_t)i * m + q] * y[q];
    y[i] = s / H[(size_t)i * m + i];
  }
  return y;
}

// Restarted GMRES(m) on device vectors: x (initial guess in, solution out)
// for A x = b with the plan's operator (fp64: the all-FP64 one). Stops when
// |g(k)| / ||b - A x0|| < tol (opGMRES.hpp:207) or after maxit iterations.
// Returns the iteration count; *est the final estimate.
int gmres_core(Solver& S, Krylov& K, bool fp64, const zd* b, zd* x, int m, double tol, int maxit, double* est_out,
               double* hist, int hist_cap, bool* conv) {
  using C = std::complex<double>;
  const long long n = S.n;
  S.residual(b, x, K.tmp.p, fp64);
  const double r0n = S.norm(K, K.tmp.p);
  double beta = r0n;
  bool converged = (r0n == 0.0);
  double est = converged ? 0.0 : 1.0;
  int it = 0;
  std::vector<C> H((size_t)(m + 1) * m), g(m + 1), hv;
  std::vector<Givens> rot(m);
  while (!converged && it < maxit) {
    // V[:,0] = r / beta (norm recomputed on the device from K.h[0])
    S.norm2(K, K.tmp.p, K.h.p);
    hipLaunchKernelGGL(k_normalize, dim3(S.nb), dim3(256), 0, S.st, n, K.tmp.p, K.h.p, K.V.p);
    std::fill(g.begin(), g.end(), C(0.0));
    g[0] = beta;
    int j = 0;
    for (; j < m && it < maxit; j++) {
      it++;
      S.apply(K.V.p + (size_t)j * n, K.w.p, fp64);
      S.arnoldi_step(K, m, j, hv);
      for (int i = 0; i < j; i++) apply_givens(rot[i], hv[i], hv[i + 1]);
      rot[j] = make_givens(hv[j], hv[j + 1]);
      apply_givens(rot[j], hv[j], hv[j + 1]);
      apply_givens(rot[j], g[j], g[j + 1]);
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = hv[i];
      est = std::abs(g[j + 1]) / r0n;
      if (hist && it - 1 < hist_cap) hist[it - 1] = est;
      if (est < tol) {
        converged = true;
        j++;
        break;
      }
    }
    S.update(K, K.V.p, j, back_substitute(H, m, g, j), x);
    if (!converged && it < maxit) {
      S.residual(b, x, K.tmp.p, fp64);
      beta = S.norm(K, K.tmp.p);
    }
  }
  *est_out = est;
  *conv = converged;
  return it;
}

}  // namespace

extern "C" int ifgf_gmres(ifgf_plan* P, const double* rhs, double* x, int restart, double tol, int maxit,
                          int* iters, double* relres, double* hist, int hist_cap) {
  if (!P || !rhs || !x || !iters || !relres || restart < 1 || maxit < 1)
    return plan_fail(P, IFGF_ERR_ARG, "invalid argument");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long n = P->geo.N;
  *iters = 0;
  try {
    Solver S(P);
    Krylov K(n, restart);
    DevBuf<zd> b, xd;
    b.alloc(n);
    xd.alloc(n);
    HIP_TRY(hipMemcpyAsync(b.p, rhs, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    HIP_TRY(hipMemcpyAsync(xd.p, x, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    const double bnorm = S.norm(K, b.p);
    if (bnorm == 0.0) {
      std::fill((std::complex<double>*)x, (std::complex<double>*)x + n, std::complex<double>(0.0));
      *relres = 0.0;
      return IFGF_OK;
    }
    double est = 0.0;
    bool converged = false;
    *iters = gmres_core(S, K, false, b.p, xd.p, restart, tol, maxit, &est, hist, hist_cap, &converged);
    // true residual with the all-FP64 operator
    S.residual(b.p, xd.p, K.tmp.p, true);
    *relres = S.norm(K, K.tmp.p) / bnorm;
    HIP_TRY(hipMemcpyAsync(x, xd.p, n * sizeof(zd), hipMemcpyDeviceToHost, S.st));
    HIP_TRY(hipStreamSynchronize(S.st));
    return converged ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}

extern "C" int ifgf_fgmres(ifgf_plan* P, const double* rhs, double* x, int restart, double tol, int maxit,
                           int inner_restart, double inner_floor, int* outer_iters, int* inner_iters,
                           double* relres, double* hist, int hist_cap) {
  if (!P || !rhs || !x || !outer_iters || !inner_iters || !relres || restart < 1 || maxit < 1 || inner_restart < 1 ||
      !(inner_floor > 0.0))
    return plan_fail(P, IFGF_ERR_ARG, "invalid argument");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  using C = std::complex<double>;
  const long long n = P->geo.N;
  const int m = restart;
  *outer_iters = 0;
  *inner_iters = 0;
  try {
    Solver S(P);
    Krylov K(n, m), Ki(n, inner_restart);
    DevBuf<zd> b, xd, Z;
    b.alloc(n);
    xd.alloc(n);
    Z.alloc((size_t)n * m);
    HIP_TRY(hipMemcpyAsync(b.p, rhs, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    HIP_TRY(hipMemcpyAsync(xd.p, x, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    const double bnorm = S.norm(K, b.p);
    if (bnorm == 0.0) {
      std::fill((C*)x, (C*)x + n, C(0.0));
      *relres = 0.0;
      return IFGF_OK;
    }
    S.residual(b.p, xd.p, K.tmp.p, true);
    const double r0n = S.norm(K, K.tmp.p);
    double beta = r0n;
    bool converged = (r0n == 0.0);
    int it = 0, inner_total = 0;
    std::vector<C> H((size_t)(m + 1) * m), g(m + 1), hv;
    std::vector<Givens> rot(m);
    while (!converged && it < maxit) {
      S.norm2(K, K.tmp.p, K.h.p);
      hipLaunchKernelGGL(k_normalize, dim3(S.nb), dim3(256), 0, S.st, n, K.tmp.p, K.h.p, K.V.p);
      std::fill(g.begin(), g.end(), C(0.0));
      g[0] = beta;
      int j = 0;
      for (; j < m && it < maxit; j++) {
        it++;
        // z_j ~ A_mixed^-1 v_j: the outer residual must still fall by
        // tol / (current relative residual), so the inner solve needs no more
        // relative accuracy than that (never below inner_floor)
        const double cur = std::abs(g[j]) / r0n;
        const double tau = std::min(0.5, std::max(inner_floor, tol / std::max(cur, 1e-300)));
        zd* zj = Z.p + (size_t)j * n;
        HIP_TRY(hipMemsetAsync(zj, 0, n * sizeof(zd), S.st));
        double iest = 0.0;
        bool iconv = false;
        inner_total += gmres_core(S, Ki, false, K.V.p + (size_t)j * n, zj, inner_restart, tau, inner_restart, &iest,
                                  nullptr, 0, &iconv);
        // w = A_fp64 z_j, orthogonalised against V (FP64 Arnoldi)
        S.apply(zj, K.w.p, true);
        S.arnoldi_step(K, m, j, hv);
        for (int i = 0; i < j; i++) apply_givens(rot[i], hv[i], hv[i + 1]);
        rot[j] = make_givens(hv[j], hv[j + 1]);
        apply_givens(rot[j], hv[j], hv[j + 1]);
        apply_givens(rot[j], g[j], g[j + 1]);
        for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = hv[i];
        const double est = std::abs(g[j + 1]) / r0n;
        if (hist && it - 1 < hist_cap) hist[it - 1] = est;
        if (est < tol) {
          converged = true;
          j++;
          break;
        }
      }
      // x += Z y (flexible: the update uses the preconditioned directions)
      S.update(K, Z.p, j, back_substitute(H, m, g, j), xd.p);
      if (!converged && it < maxit) {
        S.residual(b.p, xd.p, K.tmp.p, true);
        beta = S.norm(K, K.tmp.p);
      }
    }
    S.residual(b.p, xd.p, K.tmp.p, true);
    *relres = S.norm(K, K.tmp.p) / bnorm;
    *outer_iters = it;
    *inner_iters = inner_total;
    HIP_TRY(hipMemcpyAsync(x, xd.p, n * sizeof(zd), hipMemcpyDeviceToHost, S.st));
    HIP_TRY(hipStreamSynchronize(S.st));
    return converged ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}
