// The near-singular search (determine_patch_near_singular_points,
// ForwardMap.tpp:224-285, one thread per candidate point) and the
// rectangular-polar singular / near-singular correction built on the
// GPU (the precomputation of ForwardMap::compute_precomputations,
// ForwardMap.tpp:391-423, per column single_patch_point_compute_precomputations,
// ForwardMap.tpp:287-389): one thread per column evaluates the 40-node RP
// rule (kernel values through the device Hankel functions), accumulates the
// Np Chebyshev moments and folds them with the Chebyshev-coefficient map, with
// the same column code (rp.h) as the host setup. The host then assembles the
// same target-major CSR.
#include <hip/hip_runtime.h>

#include <vector>

#include "plan_internal.h"
#include "rp.h"

using namespace ifgf;

namespace {

constexpr int kMaxNpDevice = 16;

__global__ void __launch_bounds__(256)
    k_rp_columns(long long ncol, const RpColumn* __restrict__ cols, const double* __restrict__ t1,
                 const double* __restrict__ t2, Curve c, int F, double k, double eta, double p, int Np,
                 const double* __restrict__ fn, const double* __restrict__ fw, int nq, const double* __restrict__ Tf,
                 zd* __restrict__ W) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  zd M[kMaxNpDevice];
  const RpColumn rc = cols[col];
  rp_column(c, F, k, eta, p, t1[rc.patch], t2[rc.patch], rc.tsing, rc.ttgt, Np, fn, fw, nq, M);
  for (int kk = 0; kk < Np; kk++) {
    zd s{0.0, 0.0};
    for (int n = 0; n < Np; n++) s = cadd(s, cscale(M[n], Tf[n * Np + kk]));
    W[(size_t)col * Np + kk] = s;
  }
}

// One thread per (patch p, adjacent patch e, node j) candidate, in the host
// loop's order: ts_out = the closest parameter, or NaN when the point is not
// near-singular to the patch.
__global__ void __launch_bounds__(256)
    k_near_singular(int npatches, int Np, const double* __restrict__ t1, const double* __restrict__ t2,
                    const double* __restrict__ x, const double* __restrict__ y, Curve c, double delta,
                    double* __restrict__ ts_out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)npatches * 2 * Np) return;
  const int j = (int)(i % Np);
  const int pe = (int)(i / Np);
  const int p = pe / 2, e = pe % 2;
  const long long q = (long long)near_singular_candidate_patch(p, e, npatches) * Np + j;
  double ts = 0.0;
  ts_out[i] = near_singular_param(c, t1[p], t2[p], x[q], y[q], delta, ts) ? ts : NAN;
}

}  // namespace

bool ifgf::near_singular_device(const Curve& c, double delta, Geometry& g, std::string& err) {
  try {
    const int P = g.npatches, Np = g.Np;
    const long long n = (long long)P * 2 * Np;
    DevBuf<double> dt1, dt2, dx, dy, dts;
    dt1.upload(g.t1);
    dt2.upload(g.t2);
    dx.upload(g.x);
    dy.upload(g.y);
    dts.alloc(n);
    hipLaunchKernelGGL(k_near_singular, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, P, Np, dt1.p, dt2.p,
                       dx.p, dy.p, c, delta, dts.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> ts(n);
    HIP_TRY(hipMemcpy(ts.data(), dts.p, n * sizeof(double), hipMemcpyDeviceToHost));
    for (int p = 0; p < P; p++)
      for (int e = 0; e < 2; e++) {
        const long long base = (long long)near_singular_candidate_patch(p, e, P) * Np;
        for (int j = 0; j < Np; j++) {
          const double t = ts[((long long)p * 2 + e) * Np + j];
          if (t != t) continue;  // NaN: not near-singular
          g.ns_idx[p].push_back(base + j);
          g.ns_t[p].push_back(t);
        }
      }
  } catch (std::exception& e) {
    err = e.what();
    return false;
  }
  return true;
}

bool ifgf::build_correction_device(const Geometry& g, const Curve& c, int F, double k, double eta, double p,
                                   Correction& out, std::string& err, hipStream_t st) {
  const int Np = g.Np;
  if (Np > kMaxNpDevice) {
    err = "device correction supports Np <= 16";
    return false;
  }
  std::vector<double> fn, fw;
  rp_rule(fn, fw);
  std::vector<RpColumn> cols;
  std::vector<long long> off;
  rp_columns(g, cols, off);
  const long long ncol = (long long)cols.size();
  DevBuf<RpColumn> dcols;
  DevBuf<double> dt1, dt2, dfn, dfw, dTf;
  DevBuf<zd> dW;
  dcols.upload(cols);
  dt1.upload(g.t1);
  dt2.upload(g.t2);
  dfn.upload(fn);
  dfw.upload(fw);
  dTf.upload(rp_fold_table(g));
  dW.alloc((size_t)ncol * Np);
  hipLaunchKernelGGL(k_rp_columns, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, st, ncol, dcols.p, dt1.p,
                     dt2.p, c, F, k, eta, p, Np, dfn.p, dfw.p, (int)fn.size(), dTf.p, dW.p);
  HIP_TRY(hipGetLastError());
  std::vector<zd> W((size_t)ncol * Np);
  HIP_TRY(hipMemcpyAsync(W.data(), dW.p, W.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  assemble_correction(g, W, off, out);
  return true;
}
