// The rectangular-polar singular / near-singular correction built on the
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

}  // namespace

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
