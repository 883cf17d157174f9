// Restarted GMRES on the GPU for the plan's forward map (ifgf_gmres).
//
// Replaces ForwardMap::solve's Eigen::opGMRES<LinearOperator>
// (ForwardMap.tpp:658-697; Householder Arnoldi in opGMRES.hpp:68-269) with an
// Arnoldi process that keeps the Krylov basis V (n x (m+1) complex, column
// major) in HBM and orthogonalises by classical Gram-Schmidt with selective
// re-orthogonalisation, all in FP64: a second Gram-Schmidt pass runs when the
// first left less than a tenth of the new vector's norm (the "twice is
// enough" test of Daniel, Gragg, Kaufman and Stewart, 1976, with eta = 0.1),
// which the device decides from the norms the first pass already summed. The
// projections stay on the device; the normalising kernel writes the new
// Hessenberg column h[0..j+1] straight into pinned host memory, where the
// Givens rotations and the residual estimate |w(k)| / ||b - A x0||
// (opGMRES.hpp:207) are updated. An iteration is four kernels besides the
// operator, six when the second pass runs (each projection's partial sums are
// reduced by the kernel that consumes them), and the host applies the
// rotations of iteration j while the device already runs iteration j+1
// (which needs only the new basis vector): at convergence that one
// speculative iteration is discarded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <functional>
#include <memory>
#include <vector>

#include "plan_internal.h"

using namespace ifgf;

namespace {

constexpr int kDot = 64;       // partial sums per projection
constexpr int kMaxRestart = 256;

// Sum of the nb values x[0..nb) in a fixed order, by the first wave of the
// block (lane l takes l, l + 64, ...; then a butterfly): every block that
// calls it gets the same bits. All threads of the block must call it.
__device__ inline double block_sum_fixed(const double* __restrict__ x, int nb) {
  __shared__ double r;
  if (threadIdx.x < 64) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += 64) a += x[i];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (threadIdx.x == 0) r = a;
  }
  __syncthreads();
  return r;
}

// part[col * kDot + blockIdx.x] = sum over this block's rows of conj(V[:,col]) * w;
// the column col == wcol (if any) takes w itself (its squared norm).
// gate (second Gram-Schmidt pass only): runs only when gate[0] is set.
__global__ void __launch_bounds__(256)
    k_dots(long long n, const zd* __restrict__ V, long long ldv, const zd* __restrict__ w, zd* __restrict__ part,
           int wcol = -1, const int* __restrict__ gate = nullptr) {
  if (gate && !gate[0]) return;
  const int col = blockIdx.y;
  const zd* v = col == wcol ? w : V + (size_t)col * ldv;
  zd acc{0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
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
  if (threadIdx.x == 0) part[(size_t)col * kDot + blockIdx.x] = zd{sx[0], sy[0]};
}

// Sum of the kDot partials of column c, in a fixed order (the same in every
// block that needs it): four interleaved accumulators, then combined.
__device__ inline zd sum_partials(const zd* __restrict__ part, int c) {
  zd a[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  for (int b = 0; b < kDot; b += 4)
#pragma unroll
    for (int q = 0; q < 4; q++) a[q] = cadd(a[q], part[(size_t)c * kDot + b + q]);
  return cadd(cadd(a[0], a[1]), cadd(a[2], a[3]));
}

// Second Gram-Schmidt pass needed: ||w'||^2 (the nbn per-block partials of
// the first pass's output) < ||w||^2 / 100 (column wcol of the first pass's
// dot partials), i.e. ||w'|| < 0.1 ||w||. With eta = 1/sqrt(2) the second
// pass ran in nearly every iteration of these second-kind systems (A v keeps
// about half of v) and cost a tenth of a 64-direction star solve; at 0.1 the
// iterations and residuals are the same. Deterministic and the same in every
// block.
// Both sums are formed by the block's first wave in a fixed order (lane l
// takes partials l, l + 64, ...; then a butterfly).
__device__ inline bool second_pass(const zd* __restrict__ part_w, int wcol, const double* __restrict__ part_n,
                                   int nbn) {
  static_assert(kDot == 64, "one dot partial per lane");
  __shared__ int two;
  if (threadIdx.x < 64) {
    double after = 0.0;
    for (int i = threadIdx.x; i < nbn; i += 64) after += part_n[i];
    double before = part_w[(size_t)wcol * kDot + threadIdx.x].x;
    for (int o = 32; o > 0; o >>= 1) {
      after += __shfl_xor(after, o, 64);
      before += __shfl_xor(before, o, 64);
    }
    if (threadIdx.x == 0) two = after < 0.01 * before ? 1 : 0;
  }
  __syncthreads();
  return two != 0;
}

// Every block reduces the projections h[c] = sum_b part[c][b] (c < ncol);
// block 0 stores them into hio (or adds them, for the second Gram-Schmidt
// pass); then each thread updates its row, w -= V h, and the block writes
// the partial sum of |w_i|^2 over its rows to part_n[blockIdx.x]. A second
// pass (gate set) runs only when gate[0] is set.
__global__ void __launch_bounds__(256)
    k_project_out(long long n, int ncol, const zd* __restrict__ V, long long ldv, const zd* __restrict__ part,
                  zd* __restrict__ hio, int accumulate, zd* __restrict__ w, double* __restrict__ part_n,
                  const int* __restrict__ gate = nullptr) {
  if (gate && !gate[0]) return;
  __shared__ zd hs[kMaxRestart + 1];
  __shared__ double red[256];
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) hs[c] = sum_partials(part, c);
  __syncthreads();
  if (blockIdx.x == 0)
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) hio[c] = accumulate ? cadd(hio[c], hs[c]) : hs[c];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (i < n) {
    zd acc = w[i];
    for (int c = 0; c < ncol; c++) cfma(acc, cscale(hs[c], -1.0), V[(size_t)c * ldv + i]);
    w[i] = acc;
    sq = fma(acc.x, acc.x, acc.y * acc.y);
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part_n[blockIdx.x] = red[0];
}

// Arnoldi step end: ||w||^2 = the first pass's partials part_n1, or the
// second's (part_n2) when the second pass ran (second[0]); V[:, j+1] =
// w / ||w||; block 0 stores ||w|| into h[ncol] and writes the column
// h[0..ncol] to hcol (pinned host memory).
__global__ void __launch_bounds__(256)
    k_arnoldi_end(long long n, int ncol, const zd* __restrict__ w, const int* __restrict__ second,
                  const double* __restrict__ part_n1, const double* __restrict__ part_n2, zd* __restrict__ dst,
                  zd* __restrict__ h, zd* __restrict__ hcol) {
  const bool two = second[0] != 0;
  const double nn = sqrt(block_sum_fixed(two ? part_n2 : part_n1, gridDim.x));
  if (blockIdx.x == 0) {
    for (int c = threadIdx.x; c <= ncol; c += blockDim.x) hcol[c] = c < ncol ? h[c] : zd{nn, 0.0};
    if (threadIdx.x == 0) h[ncol] = zd{nn, 0.0};
  }
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] = cscale(w[i], 1.0 / nn);
}

// dst = src / ||src|| with ||src||^2 = sum of the kDot partials (part from
// k_dots(src, src)); block 0 stores ||src|| into *nrm.
__global__ void __launch_bounds__(256)
    k_normalize(long long n, const zd* __restrict__ src, const zd* __restrict__ part, zd* __restrict__ dst,
                zd* __restrict__ nrm) {
  __shared__ double s;
  if (threadIdx.x == 0) s = sum_partials(part, 0).x;
  __syncthreads();
  const double nn = sqrt(s);
  if (blockIdx.x == 0 && threadIdx.x == 0 && nrm) *nrm = zd{nn, 0.0};
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] = cscale(src[i], 1.0 / nn);
}

// out[0] = sum of the kDot partials of column 0 (a squared norm)
__global__ void __launch_bounds__(64) k_sum0(const zd* __restrict__ part, zd* __restrict__ out) {
  if (threadIdx.x == 0) out[0] = sum_partials(part, 0);
}

// x -= sign * sum_{j < ncol} B[:,j] y[j] (B in FP64 or, for an FP32 Krylov
// basis, FP32 widened to FP64)
template <typename VT>
__global__ void __launch_bounds__(256)
    k_update(long long n, int ncol, const VT* __restrict__ B, long long ldb, const zd* __restrict__ y, double sign,
             zd* __restrict__ x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  zd acc = x[i];
  for (int j = 0; j < ncol; j++) cfma(acc, cscale(y[j], -sign), zwiden(B[(size_t)j * ldb + i]));
  x[i] = acc;
}

// r = b - r
__global__ void __launch_bounds__(256) k_residual(long long n, const zd* __restrict__ b, zd* __restrict__ r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = csub(b[i], r[i]);
}

// ---- several systems in lockstep (ifgf_gmres_multi) ----------------------
// Launch slot a works on system sys[a]: its basis at V + sys[a] * vs
// (columns ld n), its work vector at W + a * n (the block the batched
// operator wrote), its partials at part + a * ps, its Hessenberg column at
// H + sys[a] * hs.

// second[a] = second_pass() of slot a (one block per slot): the gated
// kernels of the second pass then read one flag per system.
__global__ void __launch_bounds__(64)
    k_gate_ms(int ncol, const zd* __restrict__ part, long long ps, const double* __restrict__ part_n, long long pns,
              int nbn, int* __restrict__ second) {
  const int a = blockIdx.x;
  const bool two = second_pass(part + (size_t)a * ps, ncol, part_n + (size_t)a * pns, nbn);
  if (threadIdx.x == 0) second[a] = two ? 1 : 0;
}

// Column wcol takes W[:, a] itself; gate (second pass): runs for slot a only
// when gate[a] is set. The basis V in FP64 or FP32 (widened; sums in FP64).
template <typename VT>
__global__ void __launch_bounds__(256)
    k_dots_ms(long long n, const VT* __restrict__ V, long long vs, const zd* __restrict__ W, zd* __restrict__ part,
              long long ps, const int* __restrict__ sys, int wcol, const int* __restrict__ gate) {
  const int col = blockIdx.y, a = blockIdx.z;
  if (gate && !gate[a]) return;
  const zd* w = W + (size_t)a * n;
  const VT* v = V + (size_t)sys[a] * vs + (size_t)col * n;
  const bool self = col == wcol;
  zd acc{0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const zd y = w[i];
    const zd x = self ? y : zwiden(v[i]);
    acc.x = fma(x.x, y.x, fma(x.y, y.y, acc.x));
    acc.y = fma(x.x, y.y, fma(-x.y, y.x, acc.y));
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
  if (threadIdx.x == 0) part[(size_t)a * ps + (size_t)col * kDot + blockIdx.x] = zd{sx[0], sy[0]};
}

// k_project_out for slot a = blockIdx.y (system sys[a]); the per-block
// squared norms go to part_n + a * pns.
template <typename VT>
__global__ void __launch_bounds__(256)
    k_project_out_ms(long long n, int ncol, const VT* __restrict__ V, long long vs, const zd* __restrict__ part,
                     long long ps, zd* __restrict__ H, long long hs, int accumulate, zd* __restrict__ W,
                     const int* __restrict__ sys, double* __restrict__ part_n, long long pns,
                     const int* __restrict__ gate) {
  const int a = blockIdx.y, s = sys[a];
  if (gate && !gate[a]) return;
  __shared__ zd hsh[kMaxRestart + 1];
  __shared__ double red[256];
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) hsh[c] = sum_partials(part + (size_t)a * ps, c);
  __syncthreads();
  zd* h = H + (size_t)s * hs;
  if (blockIdx.x == 0)
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) h[c] = accumulate ? cadd(h[c], hsh[c]) : hsh[c];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (i < n) {
    const VT* v = V + (size_t)s * vs;
    zd* w = W + (size_t)a * n;
    zd acc = w[i];
    for (int c = 0; c < ncol; c++) cfma(acc, cscale(hsh[c], -1.0), zwiden(v[(size_t)c * n + i]));
    w[i] = acc;
    sq = fma(acc.x, acc.x, acc.y * acc.y);
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part_n[(size_t)a * pns + blockIdx.x] = red[0];
}

// k_arnoldi_end for slot a = blockIdx.y: V_s[:, col] = W[:, a] / ||W[:, a]||,
// H[s][col] = the norm (the second pass's when second_pass() held).
template <typename VT>
__global__ void __launch_bounds__(256)
    k_arnoldi_end_ms(long long n, int ncol, const zd* __restrict__ W, const int* __restrict__ second,
                     const double* __restrict__ part_n1, const double* __restrict__ part_n2, long long pns,
                     VT* __restrict__ V, long long vs, zd* __restrict__ H, long long hs, const int* __restrict__ sys) {
  const int a = blockIdx.y, s = sys[a];
  const bool two = second[a] != 0;
  const double nn = sqrt(block_sum_fixed((two ? part_n2 : part_n1) + (size_t)a * pns, gridDim.x));
  if (blockIdx.x == 0 && threadIdx.x == 0) H[(size_t)s * hs + ncol] = zd{nn, 0.0};
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V[(size_t)s * vs + (size_t)ncol * n + i] = zcast<VT>(cscale(W[(size_t)a * n + i], 1.0 / nn));
}

// partials of ||W[:, a]||^2
__global__ void __launch_bounds__(256)
    k_norm_parts_ms(long long n, const zd* __restrict__ W, zd* __restrict__ part, long long ps) {
  const int a = blockIdx.z;
  const zd* w = W + (size_t)a * n;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc = fma(w[i].x, w[i].x, fma(w[i].y, w[i].y, acc));
  __shared__ double sx[256];
  sx[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sx[threadIdx.x] += sx[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)a * ps + blockIdx.x] = zd{sx[0], 0.0};
}

// V[:, col] of system sys[a] = W[:, a] / ||W[:, a]||; H[sys[a]][col] = the norm
template <typename VT>
__global__ void __launch_bounds__(256)
    k_normalize_ms(long long n, const zd* __restrict__ W, const zd* __restrict__ part, long long ps,
                   VT* __restrict__ V, long long vs, int col, zd* __restrict__ H, long long hs,
                   const int* __restrict__ sys) {
  const int a = blockIdx.y, s = sys[a];
  __shared__ double nn2;
  if (threadIdx.x == 0) nn2 = sum_partials(part + (size_t)a * ps, 0).x;
  __syncthreads();
  const double nn = sqrt(nn2);
  if (blockIdx.x == 0 && threadIdx.x == 0 && H) H[(size_t)s * hs + col] = zd{nn, 0.0};
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V[(size_t)s * vs + (size_t)col * n + i] = zcast<VT>(cscale(W[(size_t)a * n + i], 1.0 / nn));
}

// dst[:, a] = src_s[:, col] for s = sys[a] (src system stride ss, column stride n)
template <typename VT>
__global__ void __launch_bounds__(256)
    k_gather_ms(long long n, const VT* __restrict__ src, long long ss, int col, const int* __restrict__ sys,
                zd* __restrict__ dst) {
  const int a = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[(size_t)a * n + i] = zwiden(src[(size_t)sys[a] * ss + (size_t)col * n + i]);
}

// W[:, a] = B_s - W[:, a] for s = sys[a]
__global__ void __launch_bounds__(256)
    k_residual_ms(long long n, const zd* __restrict__ B, const int* __restrict__ sys, zd* __restrict__ W) {
  const int a = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  W[(size_t)a * n + i] = csub(B[(size_t)sys[a] * n + i], W[(size_t)a * n + i]);
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

// Device workspace of one GMRES cycle of restart m on vectors of length n,
// and two pinned host slots (with their events) for the Hessenberg columns of
// consecutive iterations.
struct Krylov {
  DevBuf<zd> V, w, tmp, part, h;
  DevBuf<double> part_n;  // per-block squared norms of the two Gram-Schmidt passes
  DevBuf<int> gate;       // 1 when the second Gram-Schmidt pass runs
  zd* hpin = nullptr;
  zd* hpin_dev = nullptr;  // hpin as a device pointer
  hipEvent_t ev[2] = {nullptr, nullptr};
  int m;
  long long n;
  Krylov(long long n_, int m_) : m(m_), n(n_) {
    V.alloc((size_t)n * (m + 1));
    w.alloc(n);
    tmp.alloc(n);
    part.alloc((size_t)kDot * (2 * m + 4));
    h.alloc(m + 2);
    part_n.alloc((size_t)2 * ((n + 255) / 256));
    gate.alloc(1);
    HIP_TRY(hipHostMalloc((void**)&hpin, 2 * (m + 2) * sizeof(zd), hipHostMallocDefault));
    HIP_TRY(hipHostGetDevicePointer((void**)&hpin_dev, hpin, 0));
    for (auto& e : ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  ~Krylov() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (hpin) (void)hipHostFree(hpin);
  }
  Krylov(const Krylov&) = delete;
  Krylov& operator=(const Krylov&) = delete;
};

using Clock = std::chrono::steady_clock;
inline double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

struct Solver {
  ifgf_plan* P;
  hipStream_t st;
  long long n;
  int nb;
  double wait_s = 0.0;  // host time blocked on the device (event / stream synchronisations)
  Solver(ifgf_plan* p) : P(p), st(p->stream), n(p->geo.N), nb((int)((p->geo.N + 255) / 256)) {}
  void sync() {
    const auto t0 = Clock::now();
    HIP_TRY(hipStreamSynchronize(st));
    wait_s += seconds_since(t0);
  }

  void apply(const zd* x, zd* y, bool fp64) {
    const int rc = apply_device(P, x, y, st, fp64);
    if (rc) throw std::runtime_error(P->err);
  }
  // ||v|| (synchronises)
  double norm(Krylov& K, const zd* v) {
    hipLaunchKernelGGL(k_dots, dim3(kDot, 1), dim3(256), 0, st, n, v, n, v, K.part.p);
    hipLaunchKernelGGL(k_sum0, dim3(1), dim3(64), 0, st, K.part.p, K.h.p);
    zd h0;
    HIP_TRY(hipMemcpyAsync(&h0, K.h.p, sizeof(zd), hipMemcpyDeviceToHost, st));
    sync();
    return std::sqrt(h0.x);
  }
  // dst = src / ||src|| (asynchronous)
  void normalize(Krylov& K, const zd* src, zd* dst, zd* nrm) {
    hipLaunchKernelGGL(k_dots, dim3(kDot, 1), dim3(256), 0, st, n, src, n, src, K.part.p);
    hipLaunchKernelGGL(k_normalize, dim3(nb), dim3(256), 0, st, n, src, K.part.p, dst, nrm);
  }
  // r = b - A x
  void residual(const zd* b, const zd* x, zd* r, bool fp64) {
    apply(x, r, fp64);
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, st, n, b, r);
  }
  // Orthogonalises w against V[:, 0..j] (classical Gram-Schmidt, a second
  // pass when the first left less than a tenth of ||w||), normalises it
  // into V[:, j+1] and writes the Hessenberg column h[0..j+1] into pinned
  // slot j % 2, recording its event. Asynchronous. The second pass's dot
  // partials go after the first's (columns nc+1 .. 2nc), so the first pass's
  // ||w||^2 partials (column nc) stay readable for the test.
  void arnoldi_enqueue(Krylov& K, int j) {
    const int nc = j + 1;
    zd* part2 = K.part.p + (size_t)(nc + 1) * kDot;
    double* pn1 = K.part_n.p;
    double* pn2 = K.part_n.p + nb;
    hipLaunchKernelGGL(k_dots, dim3(kDot, nc + 1), dim3(256), 0, st, n, K.V.p, n, K.w.p, K.part.p, nc,
                       (const int*)nullptr);
    hipLaunchKernelGGL(k_project_out, dim3(nb), dim3(256), 0, st, n, nc, K.V.p, n, K.part.p, K.h.p, 0, K.w.p, pn1,
                       (const int*)nullptr);
    hipLaunchKernelGGL(k_gate_ms, dim3(1), dim3(64), 0, st, nc, K.part.p, 0ll, pn1, 0ll, nb, K.gate.p);
    hipLaunchKernelGGL(k_dots, dim3(kDot, nc), dim3(256), 0, st, n, K.V.p, n, K.w.p, part2, nc,
                       (const int*)K.gate.p);
    hipLaunchKernelGGL(k_project_out, dim3(nb), dim3(256), 0, st, n, nc, K.V.p, n, part2, K.h.p, 1, K.w.p, pn2,
                       (const int*)K.gate.p);
    zd* slot = K.hpin_dev + (size_t)(j % 2) * (K.m + 2);
    hipLaunchKernelGGL(k_arnoldi_end, dim3(nb), dim3(256), 0, st, n, nc, K.w.p, (const int*)K.gate.p, pn1, pn2,
                       K.V.p + (size_t)(j + 1) * n, K.h.p, slot);
    HIP_TRY(hipEventRecord(K.ev[j % 2], st));
  }
  // waits for iteration j's Hessenberg column
  void arnoldi_wait(Krylov& K, int j, std::vector<std::complex<double>>& hv) {
    const auto t0 = Clock::now();
    HIP_TRY(hipEventSynchronize(K.ev[j % 2]));
    wait_s += seconds_since(t0);
    const zd* slot = K.hpin + (size_t)(j % 2) * (K.m + 2);
    hv.resize(j + 2);
    for (int i = 0; i < j + 2; i++) hv[i] = std::complex<double>(slot[i].x, slot[i].y);
  }
  // x += B y for the first kk columns of B (ld n)
  void update(Krylov& K, const zd* B, int kk, const std::vector<std::complex<double>>& y, zd* x) {
    HIP_TRY(hipMemcpyAsync(K.h.p, y.data(), kk * sizeof(zd), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_update<zd>, dim3(nb), dim3(256), 0, st, n, kk, B, n, K.h.p, -1.0, x);
  }
};

}  // namespace

// Device workspaces of the single-system solvers, kept by the plan between
// calls so that a warm solve allocates nothing (hipMalloc / hipHostMalloc of
// tens of MB per call cost milliseconds and varied from run to run): the
// outer and inner Krylov spaces (reused while their restart is at least the
// one asked for), the right-hand side, the iterate and FGMRES's
// preconditioned directions Z.
struct ifgf::SolverCache {
  std::unique_ptr<Krylov> outer, inner;
  DevBuf<zd> b, x, Z;
  int zcols = 0;
};

void ifgf::destroy_solver_cache(SolverCache* c) { delete c; }

namespace {

Krylov& cached_krylov(std::unique_ptr<Krylov>& k, long long n, int m) {
  if (!k || k->n != n || k->m < m) {
    k.reset();
    k.reset(new Krylov(n, m));
  }
  return *k;
}

SolverCache& solver_cache(ifgf_plan* P) {
  if (!P->solver) P->solver = new SolverCache();
  SolverCache& c = *P->solver;
  const long long n = P->geo.N;
  if ((long long)c.b.n != n) {
    c.b.alloc(n);
    c.x.alloc(n);
  }
  return c;
}

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
// Returns the iteration count; *est the final estimate. x0_zero: x is zero on
// entry, so r0 = b without an application.
// ref_norm > 0: the estimates are relative to ref_norm instead of ||b - A x0||.
int gmres_core(Solver& S, Krylov& K, bool fp64, const zd* b, zd* x, int m, double tol, int maxit, double* est_out,
               double* hist, int hist_cap, bool* conv, bool x0_zero = false, double ref_norm = 0.0) {
  using C = std::complex<double>;
  const long long n = S.n;
  if (x0_zero) HIP_TRY(hipMemcpyAsync(K.tmp.p, b, n * sizeof(zd), hipMemcpyDeviceToDevice, S.st));  // r0 = b
  else S.residual(b, x, K.tmp.p, fp64);
  const double r0 = S.norm(K, K.tmp.p);
  const double r0n = ref_norm > 0.0 ? ref_norm : r0;
  double beta = r0;
  bool converged = (r0 == 0.0) || r0 / r0n < tol;
  double est = r0 == 0.0 ? 0.0 : r0 / r0n;
  int it = 0;
  std::vector<C> H((size_t)(m + 1) * m), g(m + 1), hv;
  std::vector<Givens> rot(m);
  auto enqueue = [&](int j) {
    S.apply(K.V.p + (size_t)j * n, K.w.p, fp64);
    S.arnoldi_enqueue(K, j);
  };
  while (!converged && it < maxit) {
    S.normalize(K, K.tmp.p, K.V.p, nullptr);  // V[:,0] = r / ||r||
    std::fill(g.begin(), g.end(), C(0.0));
    g[0] = beta;
    const int jmax = std::min(m, maxit - it);
    enqueue(0);
    int queued = 1;  // iterations [0, queued) are enqueued
    double e1 = 0.0, e2 = 0.0;  // the last two residual estimates of this cycle
    int j = 0;
    for (; j < jmax; j++) {
      // iteration j + 1 needs only V[:, j+1]: enqueue it before waiting for
      // j, unless the estimates' last ratio predicts that j converges (then
      // it would be thrown away)
      const bool likely_last = j >= 2 && e2 > 0.0 && e1 * (e1 / e2) < 1.5 * tol;
      if (queued == j + 1 && j + 1 < jmax && !likely_last) enqueue(queued++);
      S.arnoldi_wait(K, j, hv);
      hv[j + 1] = std::sqrt(std::norm(hv[j + 1]));
      for (int i = 0; i < j; i++) apply_givens(rot[i], hv[i], hv[i + 1]);
      rot[j] = make_givens(hv[j], hv[j + 1]);
      apply_givens(rot[j], hv[j], hv[j + 1]);
      apply_givens(rot[j], g[j], g[j + 1]);
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = hv[i];
      it++;
      est = std::abs(g[j + 1]) / r0n;
      if (hist && it - 1 < hist_cap) hist[it - 1] = est;
      if (est < tol) {
        converged = true;
        j++;
        break;
      }
      e2 = e1;
      e1 = est;
      if (queued == j + 1 && j + 1 < jmax) enqueue(queued++);  // the prediction missed
    }
    // x += V y (after the speculative iteration, if one is in flight: it only
    // wrote w, V[:, j+1] and the Hessenberg buffers)
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
  if (restart > kMaxRestart) return plan_fail(P, IFGF_ERR_ARG, "restart above 256");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long n = P->geo.N;
  *iters = 0;
  const auto t_start = Clock::now();
  try {
    Solver S(P);
    SolverCache& W = solver_cache(P);
    Krylov& K = cached_krylov(W.outer, n, restart);
    DevBuf<zd>& b = W.b;
    DevBuf<zd>& xd = W.x;
    P->solver_s[0] = seconds_since(t_start);
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
    bool zero = true;
    for (long long i = 0; i < 2 * n && zero; i++) zero = x[i] == 0.0;
    *iters = gmres_core(S, K, false, b.p, xd.p, restart, tol, maxit, &est, hist, hist_cap, &converged, zero);
    // true residual with the all-FP64 operator
    S.residual(b.p, xd.p, K.tmp.p, true);
    *relres = S.norm(K, K.tmp.p) / bnorm;
    HIP_TRY(hipMemcpyAsync(x, xd.p, n * sizeof(zd), hipMemcpyDeviceToHost, S.st));
    S.sync();
    P->solver_s[1] = S.wait_s;
    P->solver_s[2] = 0.0;
    P->solver_s[3] = seconds_since(t_start);
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
  if (restart > kMaxRestart || inner_restart > kMaxRestart) return plan_fail(P, IFGF_ERR_ARG, "restart above 256");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  using C = std::complex<double>;
  const long long n = P->geo.N;
  const int m = restart;
  *outer_iters = 0;
  *inner_iters = 0;
  const auto t_start = Clock::now();
  try {
    Solver S(P);
    SolverCache& W = solver_cache(P);
    Krylov& K = cached_krylov(W.outer, n, m);
    Krylov& Ki = cached_krylov(W.inner, n, inner_restart);
    DevBuf<zd>& b = W.b;
    DevBuf<zd>& xd = W.x;
    if (W.zcols < m || (long long)W.Z.n != (long long)n * W.zcols) {
      W.Z.alloc((size_t)n * m);
      W.zcols = m;
    }
    DevBuf<zd>& Z = W.Z;
    P->solver_s[0] = seconds_since(t_start);
    double inner_s = 0.0;
    HIP_TRY(hipMemcpyAsync(b.p, rhs, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    HIP_TRY(hipMemcpyAsync(xd.p, x, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    const double bnorm = S.norm(K, b.p);
    if (bnorm == 0.0) {
      std::fill((C*)x, (C*)x + n, C(0.0));
      *relres = 0.0;
      return IFGF_OK;
    }
    bool zero = true;
    for (long long i = 0; i < 2 * n && zero; i++) zero = x[i] == 0.0;
    if (zero) HIP_TRY(hipMemcpyAsync(K.tmp.p, b.p, n * sizeof(zd), hipMemcpyDeviceToDevice, S.st));  // r0 = b
    else S.residual(b.p, xd.p, K.tmp.p, true);
    const double r0n = S.norm(K, K.tmp.p);
    double beta = r0n;
    bool converged = (r0n == 0.0);
    int it = 0, inner_total = 0;
    std::vector<C> H((size_t)(m + 1) * m), g(m + 1), hv;
    std::vector<Givens> rot(m);
    while (!converged && it < maxit) {
      S.normalize(K, K.tmp.p, K.V.p, nullptr);
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
        const auto t_in = Clock::now();
        inner_total += gmres_core(S, Ki, false, K.V.p + (size_t)j * n, zj, inner_restart, tau, inner_restart, &iest,
                                  nullptr, 0, &iconv, true);
        inner_s += seconds_since(t_in);
        // w = A_fp64 z_j, orthogonalised against V (FP64 Arnoldi)
        S.apply(zj, K.w.p, true);
        S.arnoldi_enqueue(K, j);
        S.arnoldi_wait(K, j, hv);
        hv[j + 1] = std::sqrt(std::norm(hv[j + 1]));
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
    S.sync();
    P->solver_s[1] = S.wait_s;
    P->solver_s[2] = inner_s;
    P->solver_s[3] = seconds_since(t_start);
    return converged ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}

// Restarted GMRES for nrhs right-hand sides of the same operator, advanced in
// lockstep: each iteration applies the operator to the active systems' new
// basis vectors in one call (batched column blocks), then one set of CGS2
// kernels covers all of them; a system leaves the active set when its own
// estimate |w(k)| / ||r0|| < tol (opGMRES.hpp:207). Every system follows the
// same algorithm as ifgf_gmres (same restart, same stopping rule); all
// active systems restart together.
// The lockstep GMRES on device columns: B (right-hand sides) and X (initial
// guesses in, solutions out), n x nrhs each. true_res: relres[s] = the true
// relative residual with the all-FP64 operator; else the residual norm at the
// start of the last cycle over ||b_s||. VT: the Krylov basis in FP64 (zd) or
// FP32 (zf, for the inner corrections of the mixed-precision refinement,
// whose tolerance is 1e-5 or looser and whose operator is FP32; projections,
// norms and the work vectors stay FP64).
template <typename VT>
static int gmres_multi_dev(ifgf_plan* P, const zd* Bp, zd* Xp, int nrhs, int restart, double tol, int maxit,
                           int* iters, double* relres, bool true_res, bool x0_zero = false) {
  using C = std::complex<double>;
  const long long n = P->geo.N;
  const int m = restart, R = nrhs;
  const long long vs = n * (m + 1), hs = m + 2, ps = (long long)kDot * (2 * m + 4);
  const int nb = (int)((n + 255) / 256);
  const long long pns = 2 * (long long)nb;
  try {
    hipStream_t st = P->stream;
    DevBuf<VT> V;
    DevBuf<zd> W, part, H;
    DevBuf<double> part_n;
    DevBuf<int> sysd, gate;
    V.alloc((size_t)vs * R);
    W.alloc((size_t)n * 2 * R);  // operator input block, then its output block
    part.alloc((size_t)ps * R);
    part_n.alloc((size_t)pns * R);
    gate.alloc(R);
    H.alloc((size_t)hs * R);
    sysd.alloc(R);
    std::vector<int> act(R);
    for (int s = 0; s < R; s++) act[s] = s;
    auto upload_act = [&]() { HIP_TRY(hipMemcpyAsync(sysd.p, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice, st)); };
    // norms of W[:, 0..na) (synchronises)
    auto norms = [&](int na, std::vector<double>& out) {
      hipLaunchKernelGGL(k_norm_parts_ms, dim3(kDot, 1, na), dim3(256), 0, st, n, W.p, part.p, ps);
      std::vector<zd> pt((size_t)ps * na);
      HIP_TRY(hipMemcpyAsync(pt.data(), part.p, pt.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      out.assign(na, 0.0);
      for (int a = 0; a < na; a++) {
        double acc[4] = {0, 0, 0, 0};
        for (int b = 0; b < kDot; b += 4)
          for (int q = 0; q < 4; q++) acc[q] += pt[(size_t)a * ps + b + q].x;
        out[a] = std::sqrt((acc[0] + acc[1]) + (acc[2] + acc[3]));
      }
    };
    // W[:, a] = b_s - A x_s for the active systems (x gathered into W[:, R + a])
    auto residuals = [&](bool fp64) {
      upload_act();
      const int na = (int)act.size();
      zd* in = W.p + (size_t)R * n;
      hipLaunchKernelGGL(k_gather_ms<zd>, dim3(nb, na), dim3(256), 0, st, n, (const zd*)Xp, n, 0, sysd.p, in);
      int rc = apply_columns(P, in, n, W.p, na, st, fp64);
      if (rc) throw std::runtime_error(P->err);
      hipLaunchKernelGGL(k_residual_ms, dim3(nb, na), dim3(256), 0, st, n, Bp, sysd.p, W.p);
    };
    // ||b||
    std::vector<double> bn(R), r0n(R), beta(R), tmp;
    {
      HIP_TRY(hipMemcpyAsync(W.p, Bp, (size_t)n * R * sizeof(zd), hipMemcpyDeviceToDevice, st));
      norms(R, bn);
    }
    for (int s = 0; s < R; s++) iters[s] = 0;
    if (x0_zero) {
      // A 0 = 0: the initial residuals are the right-hand sides, which W holds
      tmp = bn;
    } else {
      residuals(false);
      norms(R, tmp);
    }
    std::vector<char> done(R, 0);
    std::vector<std::vector<C>> Hm(R, std::vector<C>((size_t)(m + 1) * m)), g(R, std::vector<C>(m + 1));
    std::vector<std::vector<Givens>> rot(R, std::vector<Givens>(m));
    for (int s = 0; s < R; s++) {
      r0n[s] = beta[s] = tmp[s];
      if (tmp[s] == 0.0 || bn[s] == 0.0) done[s] = 1;
    }
    std::vector<C> y;
    std::vector<zd> hh((size_t)hs * R);
    int it = 0;
    while (it < maxit) {
      act.clear();
      for (int s = 0; s < R; s++)
        if (!done[s]) act.push_back(s);
      if (act.empty()) break;
      // W holds the active systems' residuals in act order: V_s[:, 0] = r / ||r||
      upload_act();
      int na = (int)act.size();
      hipLaunchKernelGGL(k_norm_parts_ms, dim3(kDot, 1, na), dim3(256), 0, st, n, W.p, part.p, ps);
      hipLaunchKernelGGL(k_normalize_ms<VT>, dim3(nb, na), dim3(256), 0, st, n, W.p, part.p, ps, V.p, vs, 0,
                         (zd*)nullptr, hs, sysd.p);
      for (int s : act) {
        std::fill(g[s].begin(), g[s].end(), C(0.0));
        g[s][0] = beta[s];
      }
      const int jmax = std::min(m, maxit - it);
      std::vector<int> cols(R, 0);  // Arnoldi columns used by each system this cycle
      int j = 0;
      for (; j < jmax && na > 0; j++) {
        // Wa = A V[:, j] for the active systems (inputs gathered into W[:, R + a])
        upload_act();
        zd* in = W.p + (size_t)R * n;
        zd* Wa = W.p;
        hipLaunchKernelGGL(k_gather_ms<VT>, dim3(nb, na), dim3(256), 0, st, n, (const VT*)V.p, vs, j, sysd.p, in);
        int rc = apply_columns(P, in, n, Wa, na, st, false);
        if (rc) return rc;
        // Gram-Schmidt with selective re-orthogonalisation, as in arnoldi_enqueue
        const int nc = j + 1;
        zd* part2 = part.p + (size_t)(nc + 1) * kDot;
        double* pn1 = part_n.p;
        double* pn2 = part_n.p + nb;
        hipLaunchKernelGGL(k_dots_ms<VT>, dim3(kDot, nc + 1, na), dim3(256), 0, st, n, (const VT*)V.p, vs, Wa, part.p, ps, sysd.p, nc,
                           (const int*)nullptr);
        hipLaunchKernelGGL(k_project_out_ms<VT>, dim3(nb, na), dim3(256), 0, st, n, nc, (const VT*)V.p, vs, part.p, ps, H.p, hs, 0, Wa,
                           sysd.p, pn1, pns, (const int*)nullptr);
        hipLaunchKernelGGL(k_gate_ms, dim3(na), dim3(64), 0, st, nc, part.p, ps, pn1, pns, nb, gate.p);
        hipLaunchKernelGGL(k_dots_ms<VT>, dim3(kDot, nc, na), dim3(256), 0, st, n, (const VT*)V.p, vs, Wa, part2, ps, sysd.p, nc,
                           (const int*)gate.p);
        hipLaunchKernelGGL(k_project_out_ms<VT>, dim3(nb, na), dim3(256), 0, st, n, nc, (const VT*)V.p, vs, part2, ps, H.p, hs, 1, Wa,
                           sysd.p, pn2, pns, (const int*)gate.p);
        hipLaunchKernelGGL(k_arnoldi_end_ms<VT>, dim3(nb, na), dim3(256), 0, st, n, nc, Wa, (const int*)gate.p, pn1, pn2,
                           pns, V.p, vs, H.p, hs, sysd.p);
        HIP_TRY(hipMemcpyAsync(hh.data(), H.p, hh.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        it++;
        std::vector<int> still;
        for (int s : act) {
          std::vector<C> hv(j + 2);
          for (int i = 0; i < j + 2; i++) hv[i] = C(hh[(size_t)s * hs + i].x, hh[(size_t)s * hs + i].y);
          for (int i = 0; i < j; i++) apply_givens(rot[s][i], hv[i], hv[i + 1]);
          rot[s][j] = make_givens(hv[j], hv[j + 1]);
          apply_givens(rot[s][j], hv[j], hv[j + 1]);
          apply_givens(rot[s][j], g[s][j], g[s][j + 1]);
          for (int i = 0; i <= j; i++) Hm[s][(size_t)i * m + j] = hv[i];
          iters[s]++;
          cols[s] = j + 1;
          if (std::abs(g[s][j + 1]) / r0n[s] < tol) done[s] = 1;
          else still.push_back(s);
        }
        // systems that converged now: x_s += V_s y_s
        for (int s : act) {
          if (!done[s]) continue;
          y = back_substitute(Hm[s], m, g[s], cols[s]);
          HIP_TRY(hipMemcpyAsync(H.p + (size_t)s * hs, y.data(), cols[s] * sizeof(zd), hipMemcpyHostToDevice, st));
          hipLaunchKernelGGL(k_update<VT>, dim3(nb), dim3(256), 0, st, n, cols[s], (const VT*)V.p + (size_t)s * vs, n,
                             H.p + (size_t)s * hs, -1.0, Xp + (size_t)s * n);
        }
        act = still;
        na = (int)act.size();
      }
      // end of the cycle: the systems still active update x and restart
      for (int s : act) {
        y = back_substitute(Hm[s], m, g[s], cols[s]);
        HIP_TRY(hipMemcpyAsync(H.p + (size_t)s * hs, y.data(), cols[s] * sizeof(zd), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_update<VT>, dim3(nb), dim3(256), 0, st, n, cols[s], (const VT*)V.p + (size_t)s * vs, n,
                           H.p + (size_t)s * hs, -1.0, Xp + (size_t)s * n);
      }
      if (!act.empty() && it < maxit) {
        residuals(false);
        norms((int)act.size(), tmp);
        for (size_t a = 0; a < act.size(); a++) beta[act[a]] = tmp[a];
      }
    }
    if (true_res) {  // true residuals with the all-FP64 operator
      act.resize(R);
      for (int s = 0; s < R; s++) act[s] = s;
      residuals(true);
      norms(R, tmp);
      for (int s = 0; s < R; s++) relres[s] = bn[s] > 0.0 ? tmp[s] / bn[s] : 0.0;
    } else {
      for (int s = 0; s < R; s++) relres[s] = bn[s] > 0.0 ? beta[s] / bn[s] : 0.0;
    }
    HIP_TRY(hipStreamSynchronize(st));
    bool all = true;
    for (int s = 0; s < R; s++) all = all && done[s];
    return all ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}

// True when every one of the count doubles is zero (either sign; a NaN is not).
static bool host_all_zero(const double* v, long long count) {
  int nz = 0;
#pragma omp parallel for reduction(| : nz) schedule(static)
  for (long long i = 0; i < count; i++) nz |= v[i] != 0.0;
  return nz == 0;
}

extern "C" int ifgf_gmres_multi(ifgf_plan* P, const double* rhs, double* x, int nrhs, int restart, double tol,
                                int maxit, int* iters, double* relres) {
  if (!P || !rhs || !x || !iters || !relres || nrhs < 1 || restart < 1 || maxit < 1)
    return plan_fail(P, IFGF_ERR_ARG, "invalid argument");
  if (restart > kMaxRestart) return plan_fail(P, IFGF_ERR_ARG, "restart above 256");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long n = P->geo.N;
  try {
    hipStream_t st = P->stream;
    DevBuf<zd> B, X;
    B.alloc((size_t)n * nrhs);
    X.alloc((size_t)n * nrhs);
    HIP_TRY(hipMemcpyAsync(B.p, rhs, (size_t)n * nrhs * sizeof(zd), hipMemcpyHostToDevice, st));
    const bool x0_zero = host_all_zero(x, 2 * n * (long long)nrhs);
    if (x0_zero) HIP_TRY(hipMemsetAsync(X.p, 0, (size_t)n * nrhs * sizeof(zd), st));
    else HIP_TRY(hipMemcpyAsync(X.p, x, (size_t)n * nrhs * sizeof(zd), hipMemcpyHostToDevice, st));
    const int rc = gmres_multi_dev<zd>(P, B.p, X.p, nrhs, restart, tol, maxit, iters, relres, true, x0_zero);
    if (rc != IFGF_OK && rc != IFGF_ERR_NOT_CONVERGED) return rc;
    HIP_TRY(hipMemcpyAsync(x, X.p, (size_t)n * nrhs * sizeof(zd), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return rc;
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}

// X[:, sys[a]] += D[:, a]
__global__ void __launch_bounds__(256)
    k_add_ms(long long n, const zd* __restrict__ D, const int* __restrict__ sys, zd* __restrict__ X) {
  const int a = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  zd& v = X[(size_t)sys[a] * n + i];
  v = cadd(v, D[(size_t)a * n + i]);
}

// Mixed-precision refinement for many right-hand sides: the corrections come
// from the lockstep GMRES with the plan's default operator, the residuals
// from the all-FP64 operator; every vector stays on the device.
extern "C" int ifgf_refine_multi(ifgf_plan* P, const double* rhs, double* x, int nrhs, int restart, double tol,
                                 int maxit, int max_outer, double inner_floor, int* outer, int* iters,
                                 double* relres) {
  if (!P || !rhs || !x || !outer || !iters || !relres || nrhs < 1 || restart < 1 || maxit < 1 || max_outer < 1 ||
      !(tol > 0.0) || !(inner_floor > 0.0))
    return plan_fail(P, IFGF_ERR_ARG, "invalid argument");
  if (restart > kMaxRestart) return plan_fail(P, IFGF_ERR_ARG, "restart above 256");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long n = P->geo.N;
  const int R = nrhs, nb = (int)((n + 255) / 256);
  const long long ps = kDot;
  try {
    hipStream_t st = P->stream;
    DevBuf<zd> B, X, W, G, D, part;
    DevBuf<int> sysd;
    B.alloc((size_t)n * R);
    X.alloc((size_t)n * R);
    W.alloc((size_t)n * R);  // residual block of the active systems
    G.alloc((size_t)n * R);  // their x, gathered
    D.alloc((size_t)n * R);  // their corrections
    part.alloc((size_t)ps * R);
    sysd.alloc(R);
    HIP_TRY(hipMemcpyAsync(B.p, rhs, (size_t)n * R * sizeof(zd), hipMemcpyHostToDevice, st));
    bool x_zero = host_all_zero(x, 2 * n * (long long)R);
    if (x_zero) HIP_TRY(hipMemsetAsync(X.p, 0, (size_t)n * R * sizeof(zd), st));
    else HIP_TRY(hipMemcpyAsync(X.p, x, (size_t)n * R * sizeof(zd), hipMemcpyHostToDevice, st));
    std::vector<double> bn(R);
    {
      // (one column per thread, each summed in order)
      const std::complex<double>* b = reinterpret_cast<const std::complex<double>*>(rhs);
#pragma omp parallel for schedule(static)
      for (int q = 0; q < R; q++) {
        double acc = 0.0;
        for (long long i = 0; i < n; i++) acc += std::norm(b[(size_t)q * n + i]);
        bn[q] = std::sqrt(acc);
        iters[q] = 0;
      }
    }
    std::vector<int> act(R);
    for (int q = 0; q < R; q++) act[q] = q;
    // W[:, a] = b - A x (all-FP64 operator) for the systems in act; relres of
    // each (x_zero: x = 0 and act every system in order, so W = b)
    auto residuals = [&]() {
      const int na = (int)act.size();
      if (x_zero) {
        HIP_TRY(hipMemcpyAsync(W.p, B.p, (size_t)n * na * sizeof(zd), hipMemcpyDeviceToDevice, st));
        x_zero = false;
      } else {
        HIP_TRY(hipMemcpyAsync(sysd.p, act.data(), na * sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_gather_ms<zd>, dim3(nb, na), dim3(256), 0, st, n, (const zd*)X.p, n, 0, sysd.p, G.p);
        const int rc = apply_columns(P, G.p, n, W.p, na, st, true);
        if (rc) throw std::runtime_error(P->err);
        hipLaunchKernelGGL(k_residual_ms, dim3(nb, na), dim3(256), 0, st, n, B.p, sysd.p, W.p);
      }
      hipLaunchKernelGGL(k_norm_parts_ms, dim3(kDot, 1, na), dim3(256), 0, st, n, W.p, part.p, ps);
      std::vector<zd> pt((size_t)ps * na);
      HIP_TRY(hipMemcpyAsync(pt.data(), part.p, pt.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      for (int a = 0; a < na; a++) {
        double acc = 0.0;
        for (int q = 0; q < kDot; q++) acc += pt[(size_t)a * ps + q].x;
        relres[act[a]] = bn[act[a]] > 0.0 ? std::sqrt(acc) / bn[act[a]] : 0.0;
      }
    };
    residuals();
    int k = 0;
    std::vector<int> it(R);
    std::vector<double> rr(R);
    while (k < max_outer) {
      std::vector<int> next;
      for (int q : act)
        if (relres[q] >= tol) next.push_back(q);
      if (next.size() != act.size()) {  // keep the residual block in step with act
        act.swap(next);
        if (act.empty()) break;
        residuals();
      }
      const int na = (int)act.size();
      double emax = 0.0;
      for (int q : act) emax = std::max(emax, relres[q]);
      const double tau = std::min(0.5, std::max(inner_floor, tol / emax));
      HIP_TRY(hipMemsetAsync(D.p, 0, (size_t)n * na * sizeof(zd), st));
      // corrections with the plan's default operator; an FP32 basis with the FP32 operator
      // (D = 0: the initial residuals are W)
      const int rc = P->mixed
                         ? gmres_multi_dev<zf>(P, W.p, D.p, na, restart, tau, maxit, it.data(), rr.data(), false, true)
                         : gmres_multi_dev<zd>(P, W.p, D.p, na, restart, tau, maxit, it.data(), rr.data(), false, true);
      if (rc != IFGF_OK && rc != IFGF_ERR_NOT_CONVERGED) return rc;
      k++;
      HIP_TRY(hipMemcpyAsync(sysd.p, act.data(), na * sizeof(int), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_add_ms, dim3(nb, na), dim3(256), 0, st, n, D.p, sysd.p, X.p);
      for (int a = 0; a < na; a++) iters[act[a]] += it[a];
      residuals();
    }
    HIP_TRY(hipMemcpyAsync(x, X.p, (size_t)n * R * sizeof(zd), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *outer = k;
    bool all = true;
    for (int q = 0; q < R; q++) all = all && relres[q] < tol;
    return all ? IFGF_OK : plan_fail(P, IFGF_ERR_NOT_CONVERGED, "refinement did not reach tol");
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}

extern "C" int ifgf_plan_solver_times(const ifgf_plan* P, double seconds[4]) {
  if (!P || !seconds) return plan_fail(const_cast<ifgf_plan*>(P), IFGF_ERR_ARG, "null argument");
  for (int i = 0; i < 4; i++) seconds[i] = P->solver_s[i];
  return IFGF_OK;
}

// ---- learned preconditioner (ifgf_fno_gmres) -------------------------------
namespace {

// Flexible GMRES(m) with z_j = M(v_j) from the learned network, on the plan's
// default operator: x (initial guess in, solution out). The estimates are
// |w(k)| / ref_norm. Stops on est < tol, after maxit iterations, or on
// stagnation: once more than `window` estimates exist, an estimate above
// ratio x the one `window` iterations earlier (the caller then falls back to
// plain GMRES from the current x). Returns the iterations taken.
using Precond = std::function<void(const zd* v, zd* z)>;

// z = m v elementwise (the fixed multiplier field of IFGF_FNO_FIXED_MULTIPLIER)
__global__ void __launch_bounds__(256)
    k_cmul_field(long long n, const zd* __restrict__ m, const zd* __restrict__ v, zd* __restrict__ z) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = cmul(m[i], v[i]);
}

// m = z / v where |v|^2 > 1e-6 mean |v|^2 (else 1): the multiplier the
// network applied to v (z = M(v) = m v)
__global__ void __launch_bounds__(256)
    k_field_ratio(long long n, const zd* __restrict__ z, const zd* __restrict__ v, const double* __restrict__ vnorm2,
                  zd* __restrict__ m) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = v[i].x * v[i].x + v[i].y * v[i].y;
  m[i] = a > 1e-6 * vnorm2[0] / (double)n ? cdiv(z[i], v[i]) : zd{1.0, 0.0};
}

int fgmres_learned(Solver& S, Krylov& K, zd* Z, const Precond& M, const zd* b, zd* x, int m, double tol, int maxit,
                   double ref_norm, int window, double ratio, double* hist, int hist_cap, int hist_off, bool* conv,
                   bool* stalled) {
  using C = std::complex<double>;
  const long long n = S.n;
  S.residual(b, x, K.tmp.p, false);
  double beta = S.norm(K, K.tmp.p);
  bool converged = beta == 0.0 || beta / ref_norm < tol;
  *stalled = false;
  int it = 0;
  std::vector<C> H((size_t)(m + 1) * m), g(m + 1), hv;
  std::vector<Givens> rot(m);
  std::vector<double> ests;
  while (!converged && it < maxit && !*stalled) {
    S.normalize(K, K.tmp.p, K.V.p, nullptr);
    std::fill(g.begin(), g.end(), C(0.0));
    g[0] = beta;
    int j = 0;
    for (; j < m && it < maxit; j++) {
      it++;
      zd* zj = Z + (size_t)j * n;
      M(K.V.p + (size_t)j * n, zj);
      S.apply(zj, K.w.p, false);
      S.arnoldi_enqueue(K, j);
      S.arnoldi_wait(K, j, hv);
      hv[j + 1] = std::sqrt(std::norm(hv[j + 1]));
      for (int i = 0; i < j; i++) apply_givens(rot[i], hv[i], hv[i + 1]);
      rot[j] = make_givens(hv[j], hv[j + 1]);
      apply_givens(rot[j], hv[j], hv[j + 1]);
      apply_givens(rot[j], g[j], g[j + 1]);
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = hv[i];
      const double est = std::abs(g[j + 1]) / ref_norm;
      if (hist && hist_off + it - 1 < hist_cap) hist[hist_off + it - 1] = est;
      ests.push_back(est);
      if (est < tol) {
        converged = true;
        j++;
        break;
      }
      if ((int)ests.size() > window && est > ratio * ests[ests.size() - 1 - window]) {
        *stalled = true;
        j++;
        break;
      }
    }
    S.update(K, Z, j, back_substitute(H, m, g, j), x);
    if (!converged && it < maxit && !*stalled) {
      S.residual(b, x, K.tmp.p, false);
      beta = S.norm(K, K.tmp.p);
    }
  }
  *conv = converged;
  return it;
}

}  // namespace

extern "C" int ifgf_fno_gmres(ifgf_plan* P, ifgf_fno* F, const double* rhs, double* x, int restart, double tol,
                              int maxit, int mode, int window, double ratio, int* iters, int* fallback,
                              double* relres, double* hist, int hist_cap) {
  if (!P || !F || !rhs || !x || !iters || !fallback || !relres || restart < 1 || maxit < 1 || window < 1 ||
      !(ratio > 0.0))
    return plan_fail(P, IFGF_ERR_ARG, "invalid argument");
  if (restart > kMaxRestart) return plan_fail(P, IFGF_ERR_ARG, "restart above 256");
  if (!P->on_device) return plan_fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long n = P->geo.N;
  *iters = 0;
  *fallback = 0;
  const auto t_start = Clock::now();
  try {
    Solver S(P);
    SolverCache& W = solver_cache(P);
    Krylov& K = cached_krylov(W.outer, n, restart);
    if (W.zcols < restart || (long long)W.Z.n != (long long)n * W.zcols) {
      W.Z.alloc((size_t)n * restart);
      W.zcols = restart;
    }
    P->solver_s[0] = seconds_since(t_start);
    DevBuf<zd>& b = W.b;
    DevBuf<zd>& xd = W.x;
    HIP_TRY(hipMemcpyAsync(b.p, rhs, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    HIP_TRY(hipMemcpyAsync(xd.p, x, n * sizeof(zd), hipMemcpyHostToDevice, S.st));
    const double bnorm = S.norm(K, b.p);
    if (bnorm == 0.0) {
      std::fill((std::complex<double>*)x, (std::complex<double>*)x + n, std::complex<double>(0.0));
      *relres = 0.0;
      return IFGF_OK;
    }
    bool zero = true;
    for (long long i = 0; i < 2 * n && zero; i++) zero = x[i] == 0.0;
    if ((mode & IFGF_FNO_GUESS) && zero) {
      if (fno_apply_device(F, P, b.p, xd.p, S.st)) throw std::runtime_error(P->err);
      zero = false;
    }
    bool converged = false;
    int it = 0;
    if (mode & IFGF_FNO_PRECONDITION) {
      bool stalled = false;
      DevBuf<zd> mfield;
      Precond M = [&](const zd* v, zd* z) {
        if (fno_apply_device(F, P, v, z, S.st)) throw std::runtime_error(P->err);
      };
      if (mode & IFGF_FNO_FIXED_MULTIPLIER) {
        // the multiplier the network applies to b, kept for every Krylov vector
        mfield.alloc(n);
        if (fno_apply_device(F, P, b.p, K.tmp.p, S.st)) throw std::runtime_error(P->err);
        hipLaunchKernelGGL(k_dots, dim3(kDot, 1), dim3(256), 0, S.st, n, b.p, n, b.p, K.part.p, -1,
                           (const int*)nullptr);
        hipLaunchKernelGGL(k_sum0, dim3(1), dim3(64), 0, S.st, K.part.p, K.h.p);
        hipLaunchKernelGGL(k_field_ratio, dim3(S.nb), dim3(256), 0, S.st, n, K.tmp.p, b.p, (const double*)K.h.p,
                           mfield.p);
        M = [&](const zd* v, zd* z) {
          hipLaunchKernelGGL(k_cmul_field, dim3(S.nb), dim3(256), 0, S.st, n, (const zd*)mfield.p, v, z);
        };
      }
      it = fgmres_learned(S, K, W.Z.p, M, b.p, xd.p, restart, tol, maxit, bnorm, window, ratio, hist, hist_cap, 0,
                          &converged, &stalled);
      if (stalled && !converged && it < maxit) *fallback = 1;
      zero = false;  // the flexible iterations have updated x
    }
    if (!converged && it < maxit) {
      // plain GMRES (IdentityPreconditioner, opGMRES.hpp:274) from the current
      // iterate, with the same |w(k)| / ||b|| < tol rule
      double est = 0.0;
      it += gmres_core(S, K, false, b.p, xd.p, restart, tol, maxit - it, &est, hist ? hist + it : nullptr,
                       hist ? std::max(0, hist_cap - it) : 0, &converged, zero, bnorm);
    }
    *iters = it;
    S.residual(b.p, xd.p, K.tmp.p, true);
    *relres = S.norm(K, K.tmp.p) / bnorm;
    HIP_TRY(hipMemcpyAsync(x, xd.p, n * sizeof(zd), hipMemcpyDeviceToHost, S.st));
    S.sync();
    P->solver_s[1] = S.wait_s;
    P->solver_s[2] = 0.0;
    P->solver_s[3] = seconds_since(t_start);
    return converged ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
}
