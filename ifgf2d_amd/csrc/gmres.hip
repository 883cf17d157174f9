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

}  // namespace

extern "C" int ifgf_gmres(ifgf_plan* P, const double* rhs, double* x, int restart, double tol, int maxit,
                          int* iters, double* relres, double* hist, int hist_cap) {
  if (!P || !rhs || !x || !iters || !relres || restart < 1 || maxit < 1)
    return plan_fail(P, IFGF_ERR_ARG, "invalid argument");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  using C = std::complex<double>;
  const long long n = P->geo.N;
  const int m = restart;
  *iters = 0;
  try {
    hipStream_t st = P->stream;
    DevBuf<zd> V, b, xd, w, part, h, tmp;
    V.alloc((size_t)n * (m + 1));
    b.alloc(n);
    xd.alloc(n);
    w.alloc(n);
    tmp.alloc(n);
    part.alloc((size_t)kChunks * (m + 2));
    h.alloc(2 * (m + 2));  // [0, m+2): projections; [m+2, 2m+4): second CGS pass
    HIP_TRY(hipMemcpyAsync(b.p, rhs, n * sizeof(zd), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(xd.p, x, n * sizeof(zd), hipMemcpyHostToDevice, st));
    const int nb = (int)((n + 255) / 256);
    auto norm2 = [&](const zd* v, zd* dst) {
      hipLaunchKernelGGL(k_mdot, dim3(kChunks, 1), dim3(256), 0, st, n, v, n, v, part.p);
      hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, part.p, dst, 0);
    };
    auto fetch = [&](const zd* src, int cnt, std::vector<C>& out) {
      out.resize(cnt);
      HIP_TRY(hipMemcpyAsync(out.data(), src, cnt * sizeof(zd), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    };
    // ||b||
    std::vector<C> hv;
    norm2(b.p, h.p);
    fetch(h.p, 1, hv);
    const double bnorm = std::sqrt(hv[0].real());
    if (bnorm == 0.0) {
      std::vector<C> z(n, 0.0);
      std::copy(z.begin(), z.end(), (C*)x);
      *relres = 0.0;
      return IFGF_OK;
    }
    // r0 = b - A x0  (into V[:,0] before normalisation)
    auto residual_into = [&](zd* r) {
      int rc = apply_device(P, xd.p, r, st);
      if (rc) throw std::runtime_error(P->err);
      hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, st, n, b.p, r);
    };
    residual_into(tmp.p);
    norm2(tmp.p, h.p);
    fetch(h.p, 1, hv);
    const double r0n = std::sqrt(hv[0].real());
    double beta = r0n;
    bool converged = (r0n == 0.0);
    double est = r0n / bnorm;
    int it = 0;
    std::vector<C> H((size_t)(m + 1) * m), g(m + 1);
    std::vector<Givens> rot(m);
    while (!converged && it < maxit) {
      // V[:,0] = r / beta
      hipLaunchKernelGGL(k_normalize, dim3(nb), dim3(256), 0, st, n, tmp.p, h.p, V.p);
      std::fill(g.begin(), g.end(), C(0.0));
      g[0] = beta;
      int j = 0;
      for (; j < m && it < maxit; j++) {
        it++;
        int rc = apply_device(P, V.p + (size_t)j * n, w.p, st);
        if (rc) return rc;
        // CGS2: h = V^H w; w -= V h; h2 = V^H w; w -= V h2; h += h2
        hipLaunchKernelGGL(k_mdot, dim3(kChunks, j + 1), dim3(256), 0, st, n, V.p, n, w.p, part.p);
        hipLaunchKernelGGL(k_reduce, dim3(j + 1), dim3(256), 0, st, part.p, h.p, 0);
        hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, n, j + 1, V.p, n, h.p, 1.0, w.p);
        hipLaunchKernelGGL(k_mdot, dim3(kChunks, j + 1), dim3(256), 0, st, n, V.p, n, w.p, part.p);
        hipLaunchKernelGGL(k_reduce, dim3(j + 1), dim3(256), 0, st, part.p, h.p + m + 2, 0);
        hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, n, j + 1, V.p, n, h.p + m + 2, 1.0, w.p);
        hipLaunchKernelGGL(k_reduce, dim3(j + 1), dim3(256), 0, st, part.p, h.p, 1);  // h += h2
        norm2(w.p, h.p + j + 1);
        hipLaunchKernelGGL(k_normalize, dim3(nb), dim3(256), 0, st, n, w.p, h.p + j + 1, V.p + (size_t)(j + 1) * n);
        fetch(h.p, j + 2, hv);
        hv[j + 1] = std::sqrt(hv[j + 1].real());
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
      // y = R^{-1} g; x += V y
      const int kk = j;
      std::vector<C> y(kk);
      for (int i = kk - 1; i >= 0; i--) {
        C s = g[i];
        for (int q = i + 1; q < kk; q++) s -= H[(size_t)i * m + q] * y[q];
        y[i] = s / H[(size_t)i * m + i];
      }
      HIP_TRY(hipMemcpyAsync(h.p, y.data(), kk * sizeof(zd), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, n, kk, V.p, n, h.p, -1.0, xd.p);
      if (!converged && it < maxit) {
        residual_into(tmp.p);
        norm2(tmp.p, h.p);
        fetch(h.p, 1, hv);
        beta = std::sqrt(hv[0].real());
      }
    }
    // true residual
    residual_into(tmp.p);
    norm2(tmp.p, h.p);
    fetch(h.p, 1, hv);
    *relres = std::sqrt(hv[0].real()) / bnorm;
    *iters = it;
    HIP_TRY(hipMemcpyAsync(x, xd.p, n * sizeof(zd), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return converged ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}
