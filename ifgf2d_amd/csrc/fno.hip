// Learned preconditioner: a 1D Fourier neural operator (FNO) on the boundary
// parameter t, trained on densities the solver itself produces, that supplies
// the initial density of a solve and acts as a flexible right preconditioner
// of FGMRES (no reference counterpart: the reference's GMRES takes an
// Eigen preconditioner object, IdentityPreconditioner by default,
// opGMRES.hpp:103,153,251,274; ifgf_fno_gmres plugs the network in there).
//
// Network (width 64, 32 Fourier modes, 4 Fourier layers; BF16 weights and
// activations into FP32-accumulating MFMA, FP32 master weights):
//   features X (per node, 11 channels padded to 32)
//   H0 = X Wlift + blift
//   P_l = IDFT(mix_l(DFT(H_l))) + H_l W_l + b_l,  H_{l+1} = GELU(P_l) (l < 3), H_4 = P_3
//   Q = GELU(H_4 Wp1 + bp1),  O = Q Wp2 + bp2 (2 of 16 padded columns used)
//   multiplier m = m_phys + O_0 + i O_1,  output z = m * v
// The DFT runs over the nodes with the parameter quadrature weights (the
// nodes are Fejer points of the patches, not uniform): Hhat_m =
// (1/2pi) sum_i wt_i h(t_i) e^{-i m t_i}, m < 32, as the product of the stored
// basis Tw (N x 64, bf16) with H; the inverse, sum_m c_m Re(G_m e^{i m t}),
// as T (N x 64) times the mixed modes (c_0 = 1, c_m = 2, Im G_0 dropped).
// mix_l multiplies every mode's 64 channels by a complex 64 x 64 matrix.
//
// Inputs of a vector v on the curve (the right-hand side -u_inc of a plane
// wave when the network supplies the initial density, a Krylov vector when it
// preconditions): x, y (centred, scaled), the normal, the curvature, the
// arclength over the perimeter, k L / 2000, |v| / rms(v), the local
// tangential wavenumber w = conj(v) dv/ds / (i k (|v|^2 + eps rms^2)) (real
// part: d.tau for a plane wave of direction d; its derivative per patch by
// Chebyshev differentiation on the Fejer nodes), and e.n with e the direction
// that fits Re w = e.tau best over the curve (least squares; e = d for a plane
// wave, so e.n < 0 marks the lit side and > 0 the shadow). The network outputs a
// complex multiplier field, so z is linear in v for fixed features and
// equivariant under v -> a v (a complex); m_phys = 2c / (1 + c), c =
// sqrt(1 - Re(w)^2), is the inverse symbol of the combined-field operator
// (1/2 + K - i k S) for a plane wave on a flat boundary, which the network
// corrects for curvature, shadow and the coupling along the curve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "plan_internal.h"

using namespace ifgf;

namespace {

constexpr int kW = 64;       // width
constexpr int kModes = 32;   // Fourier modes
constexpr int kS = 2 * kModes;  // spectral rows (re, im per mode)
constexpr int kL = 4;        // Fourier layers
constexpr int kCin = 11;     // input features
constexpr int kCinPad = 32;  // padded to the MFMA K step
constexpr int kOutPad = 16;  // output columns (2 used)

typedef unsigned short bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__host__ __device__ inline bf16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__device__ inline bf16_t to_bf(float v) { return f2bf(v); }
__device__ inline bf16_t to_bf(bf16_t v) { return v; }

// ---------------------------------------------------------------- layout ---
// Flat FP32 parameter vector (the layout of ifgf_fno_get_params /
// ifgf_fno_set_params): Wlift [32][64], blift [64]; per layer l: R_l
// [32 modes][64 in][64 out][re, im], W_l [64][64], b_l [64]; Wp1 [64][64],
// bp1 [64]; Wp2 [64][16], bp2 [16]. Weights are [in][out].
struct Layout {
  long long wl, bl, R[kL], W[kL], b[kL], wp1, bp1, wp2, bp2, total;
  Layout() {
    long long o = 0;
    wl = o; o += kCinPad * kW;
    bl = o; o += kW;
    for (int l = 0; l < kL; l++) {
      R[l] = o; o += (long long)kModes * kW * kW * 2;
      W[l] = o; o += kW * kW;
      b[l] = o; o += kW;
    }
    wp1 = o; o += kW * kW;
    bp1 = o; o += kW;
    wp2 = o; o += kW * kOutPad;
    bp2 = o; o += kOutPad;
    total = o;
  }
};
const Layout kLay;

// ------------------------------------------------------------------ GEMM ---
// C (M x N, FP32, row stride ldc) = [C +] A B [+ bias], BF16 operands (FP32
// inputs are rounded to BF16 as they are staged), FP32 accumulation on the
// 16x16x32 BF16 MFMA. A element (m, k) is A[m * lda + k], or A[k * lda + m]
// when TRA; B element (k, n) is B[k * ldb + n], or B[n * ldb + k] when TRB.
// 64 x 64 tiles, 4 waves of 32 x 32 (2 x 2 MFMA tiles), K in steps of 32
// through LDS. Split K (gridDim.z > 1): block z sums k in
// [z * kchunk, (z + 1) * kchunk) into part[z][M][N]; k_gemm_reduce adds the
// partials in z order (deterministic).
constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

template <typename TA_, typename TB_, bool TRA, bool TRB>
__global__ void __launch_bounds__(256)
    k_gemm(int M, int N, int K, const TA_* __restrict__ A, long long lda, const TB_* __restrict__ B, long long ldb,
           float* __restrict__ C, long long ldc, int accumulate, const float* __restrict__ bias, int kchunk,
           float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) bf16_t As[BM][LDK];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[BN][LDK];
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += BK) {
    if (!TRA) {
      const int r = tid >> 2, kc = (tid & 3) * 8;
      const int m = m0 + r;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k = k0 + kc + j;
        As[r][kc + j] = (m < M && k < ke) ? to_bf(A[(long long)m * lda + k]) : (bf16_t)0;
      }
    } else {
      const int kk = tid >> 3, mc = (tid & 7) * 8;
      const int k = k0 + kk;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int m = m0 + mc + j;
        As[mc + j][kk] = (m < M && k < ke) ? to_bf(A[(long long)k * lda + m]) : (bf16_t)0;
      }
    }
    if (!TRB) {
      const int kk = tid >> 3, nc = (tid & 7) * 8;
      const int k = k0 + kk;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int n = n0 + nc + j;
        Bs[nc + j][kk] = (n < N && k < ke) ? to_bf(B[(long long)k * ldb + n]) : (bf16_t)0;
      }
    } else {
      const int r = tid >> 2, kc = (tid & 3) * 8;
      const int n = n0 + r;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k = k0 + kc + j;
        Bs[r][kc + j] = (n < N && k < ke) ? to_bf(B[(long long)n * ldb + k]) : (bf16_t)0;
      }
    }
    __syncthreads();
    const int kg = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + i * 16 + (lane & 15)][kg]);
      b[i] = *reinterpret_cast<const bf16x8*>(&Bs[wn + i * 16 + (lane & 15)][kg]);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        float v = acc[i][j][q];
        if (part) {
          part[((long long)blockIdx.z * M + m) * N + n] = v;
        } else {
          if (bias) v += bias[n];
          if (accumulate) v += C[(long long)m * ldc + n];
          C[(long long)m * ldc + n] = v;
        }
      }
}

// C = [C +] sum_z part[z] [+ bias], z in order
__global__ void __launch_bounds__(256)
    k_gemm_reduce(int M, int N, int Z, const float* __restrict__ part, float* __restrict__ C, long long ldc,
                  int accumulate, const float* __restrict__ bias) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M * N) return;
  const int m = (int)(t / N), n = (int)(t % N);
  float v = 0.f;
  for (int z = 0; z < Z; z++) v += part[(long long)z * M * N + t];
  if (bias) v += bias[n];
  if (accumulate) v += C[(long long)m * ldc + n];
  C[(long long)m * ldc + n] = v;
}

// --------------------------------------------------------- elementwise ---
__device__ inline float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__global__ void __launch_bounds__(256) k_gelu_fwd(long long n, const float* __restrict__ P, float* __restrict__ H) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) H[i] = gelu(P[i]);
}
// dP = dH * gelu'(P)
__global__ void __launch_bounds__(256)
    k_gelu_bwd(long long n, const float* __restrict__ P, const float* __restrict__ dH, float* __restrict__ dP) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dP[i] = dH[i] * gelu_d(P[i]);
}

// out[n] (+)= sum over rows of X[:, n] (M rows, N columns, row stride ldx):
// one block per column, fixed order.
__global__ void __launch_bounds__(256)
    k_colsum(long long M, int N, const float* __restrict__ X, long long ldx, float* __restrict__ out, int accumulate) {
  const int n = blockIdx.x;
  if (n >= N) return;
  __shared__ float red[256];
  float a = 0.f;
  for (long long m = threadIdx.x; m < M; m += 256) a += X[m * ldx + n];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = (accumulate ? out[n] : 0.f) + red[0];
}

// Mode mixing, forward: for sample b, mode m, output channel o:
//   g = sum_i (a_i + i b_i) R_m[i][o],  a = Hhat[2m], b = Hhat[2m+1];
//   G[2m][o] = c_m Re g, G[2m+1][o] = c_m Im g (0 for m = 0).
__global__ void __launch_bounds__(256)
    k_mix_fwd(int B, const float* __restrict__ Hh, const bf16_t* __restrict__ R, float* __restrict__ G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kModes * kW) return;
  const int o = t % kW, m = (t / kW) % kModes, b = t / (kW * kModes);
  const float* a = Hh + ((long long)b * kS + 2 * m) * kW;
  const float* bi = a + kW;
  const bf16_t* Rm = R + (long long)m * kW * kW * 2;
  float gr = 0.f, gi = 0.f;
  for (int i = 0; i < kW; i++) {
    const float rr = bf2f(Rm[(i * kW + o) * 2]), ri = bf2f(Rm[(i * kW + o) * 2 + 1]);
    gr = fmaf(a[i], rr, fmaf(-bi[i], ri, gr));
    gi = fmaf(a[i], ri, fmaf(bi[i], rr, gi));
  }
  const float c = m == 0 ? 1.f : 2.f;
  G[((long long)b * kS + 2 * m) * kW + o] = c * gr;
  G[((long long)b * kS + 2 * m + 1) * kW + o] = m == 0 ? 0.f : c * gi;
}

// Mode mixing, backward to the input: with dg = c_m (dG[2m], dG[2m+1]) (the
// imaginary part 0 for m = 0): dHhat[2m][i] = sum_o Rr dgr + Ri dgi,
// dHhat[2m+1][i] = sum_o -Ri dgr + Rr dgi.
__global__ void __launch_bounds__(256)
    k_mix_bwd_in(int B, const float* __restrict__ dG, const bf16_t* __restrict__ R, float* __restrict__ dHh) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kModes * kW) return;
  const int i = t % kW, m = (t / kW) % kModes, b = t / (kW * kModes);
  const float c = m == 0 ? 1.f : 2.f;
  const float* dgr = dG + ((long long)b * kS + 2 * m) * kW;
  const float* dgi = dgr + kW;
  const bf16_t* Rm = R + ((long long)m * kW + i) * kW * 2;
  float da = 0.f, db = 0.f;
  for (int o = 0; o < kW; o++) {
    const float rr = bf2f(Rm[o * 2]), ri = bf2f(Rm[o * 2 + 1]);
    const float gr = c * dgr[o], gi = m == 0 ? 0.f : c * dgi[o];
    da = fmaf(rr, gr, fmaf(ri, gi, da));
    db = fmaf(-ri, gr, fmaf(rr, gi, db));
  }
  dHh[((long long)b * kS + 2 * m) * kW + i] = da;
  dHh[((long long)b * kS + 2 * m + 1) * kW + i] = db;
}

// Mode mixing, backward to the weights (summed over the B samples in order,
// added to dR): dRr[m][i][o] += a_i dgr_o + b_i dgi_o,
// dRi[m][i][o] += -b_i dgr_o + a_i dgi_o.
__global__ void __launch_bounds__(256)
    k_mix_bwd_w(int B, const float* __restrict__ Hh, const float* __restrict__ dG, float* __restrict__ dR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kModes * kW * kW) return;
  const int o = t % kW, i = (t / kW) % kW, m = t / (kW * kW);
  const float c = m == 0 ? 1.f : 2.f;
  float rr = 0.f, ri = 0.f;
  for (int b = 0; b < B; b++) {
    const float a = Hh[((long long)b * kS + 2 * m) * kW + i], bb = Hh[((long long)b * kS + 2 * m + 1) * kW + i];
    const float gr = c * dG[((long long)b * kS + 2 * m) * kW + o];
    const float gi = m == 0 ? 0.f : c * dG[((long long)b * kS + 2 * m + 1) * kW + o];
    rr = fmaf(a, gr, fmaf(bb, gi, rr));
    ri = fmaf(-bb, gr, fmaf(a, gi, ri));
  }
  dR[(long long)t * 2] += rr;
  dR[(long long)t * 2 + 1] += ri;
}

// Features of a vector v on a plan's curve (patch order), see the file header.
// geo: per node 7 geometry channels; dfac: d/ds = dfac * d/dx_patch; D: the
// Np x Np Chebyshev differentiation matrix on the Fejer nodes in [-1, 1];
// stats: [0] = sum |v|^2 (from k_sumsq).
__global__ void __launch_bounds__(256)
    k_features(long long N, int Np, const float* __restrict__ geo, const double* __restrict__ dfac,
               const double* __restrict__ D, const zd* __restrict__ v, const double* __restrict__ stats, double k,
               float* __restrict__ X, float* __restrict__ mphys) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const long long p0 = (i / Np) * Np;
  const int j = (int)(i - p0);
  zd dv{0.0, 0.0};
  for (int q = 0; q < Np; q++) {
    const double d = D[j * Np + q];
    dv.x = fma(d, v[p0 + q].x, dv.x);
    dv.y = fma(d, v[p0 + q].y, dv.y);
  }
  dv.x *= dfac[i];
  dv.y *= dfac[i];
  const double ms = stats[0] / (double)N;  // mean |v|^2
  const double rms = sqrt(ms > 0.0 ? ms : 1.0);
  const zd vi = v[i];
  const double a2 = vi.x * vi.x + vi.y * vi.y;
  // w = conj(v) dv / (i k (|v|^2 + eps ms)); conj(v) dv = (vx dx + vy dy) + i (vx dy - vy dx)
  const double pr = vi.x * dv.x + vi.y * dv.y, pi = vi.x * dv.y - vi.y * dv.x;
  const double den = k * (a2 + 1e-2 * ms);
  double wr = pi / den, wi = -pr / den;  // (pr + i pi) / (i den)
  wr = fmin(2.0, fmax(-2.0, wr));
  wi = fmin(2.0, fmax(-2.0, wi));
  float* x = X + i * kCinPad;
#pragma unroll
  for (int c = 0; c < 7; c++) x[c] = geo[i * 7 + c];
  x[7] = (float)fmin(4.0, sqrt(a2) / rms);
  x[8] = (float)wr;
  x[9] = (float)wi;
  for (int c = kCin; c < kCinPad; c++) x[c] = 0.f;
  const double cc = sqrt(fmax(0.0, 1.0 - wr * wr));
  mphys[i] = (float)(2.0 * cc / (1.0 + cc));
}

// Least-squares direction e of Re w = e.tau over the nodes (tau = (-n_y,
// n_x)): one block sums the normal equations in a fixed order, then every
// node gets X[i][10] = e.n (|e| capped at 1).
__global__ void __launch_bounds__(256)
    k_direction_sums(long long N, const float* __restrict__ geo, const float* __restrict__ X,
                     double* __restrict__ out) {
  __shared__ double red[5][256];
  double a[5] = {0, 0, 0, 0, 0};
  for (long long i = threadIdx.x; i < N; i += 256) {
    const double tx = -geo[i * 7 + 3], ty = geo[i * 7 + 2];
    const double w = X[i * kCinPad + 8];
    a[0] += tx * tx;
    a[1] += tx * ty;
    a[2] += ty * ty;
    a[3] += w * tx;
    a[4] += w * ty;
  }
  for (int q = 0; q < 5; q++) red[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int q = 0; q < 5; q++) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double det = red[0][0] * red[2][0] - red[1][0] * red[1][0];
    double ex = 0.0, ey = 0.0;
    if (det > 1e-12 * (red[0][0] + red[2][0]) * (red[0][0] + red[2][0])) {
      ex = (red[2][0] * red[3][0] - red[1][0] * red[4][0]) / det;
      ey = (red[0][0] * red[4][0] - red[1][0] * red[3][0]) / det;
    }
    const double e = sqrt(ex * ex + ey * ey);
    if (e > 1.0) {
      ex /= e;
      ey /= e;
    }
    out[1] = ex;
    out[2] = ey;
  }
}

__global__ void __launch_bounds__(256)
    k_direction_feature(long long N, const float* __restrict__ geo, const double* __restrict__ e, float* __restrict__ X) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) X[i * kCinPad + 10] = (float)(e[1] * geo[i * 7 + 2] + e[2] * geo[i * 7 + 3]);
}

// out[0] = sum |v|^2 (one block, fixed order)
__global__ void __launch_bounds__(256) k_sumsq(long long N, const zd* __restrict__ v, double* __restrict__ out) {
  __shared__ double red[256];
  double a = 0.0;
  for (long long i = threadIdx.x; i < N; i += 256) a += v[i].x * v[i].x + v[i].y * v[i].y;
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// z = (m_phys + O0 + i O1) v
__global__ void __launch_bounds__(256)
    k_output(long long N, const float* __restrict__ O, const float* __restrict__ mphys, const zd* __restrict__ v,
             zd* __restrict__ z) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double mr = (double)mphys[i] + (double)O[i * kOutPad], mi = (double)O[i * kOutPad + 1];
  const zd a = v[i];
  z[i] = zd{mr * a.x - mi * a.y, mr * a.y + mi * a.x};
}

// Loss terms of one sample: r = z - target; lt[i] = |r|^2 * inv; dO[i][0] =
// scale * 2 Re(conj(r) v) * inv, dO[i][1] = -scale * 2 Im(conj(r) v) * inv.
__global__ void __launch_bounds__(256)
    k_loss(long long N, const float* __restrict__ O, const float* __restrict__ mphys, const zd* __restrict__ v,
           const zd* __restrict__ target, double inv, double scale, float* __restrict__ dO, double* __restrict__ lt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double mr = (double)mphys[i] + (double)O[i * kOutPad], mi = (double)O[i * kOutPad + 1];
  const zd a = v[i];
  const zd r{mr * a.x - mi * a.y - target[i].x, mr * a.y + mi * a.x - target[i].y};
  lt[i] = (r.x * r.x + r.y * r.y) * inv;
  // conj(r) v = (rx ax + ry ay) + i (rx ay - ry ax)
  const double cr = r.x * a.x + r.y * a.y, ci = r.x * a.y - r.y * a.x;
  float* d = dO + i * kOutPad;
  d[0] = (float)(scale * 2.0 * cr * inv);
  d[1] = (float)(-scale * 2.0 * ci * inv);
  for (int c = 2; c < kOutPad; c++) d[c] = 0.f;
}

// out[slot] = sum x[0..n) (one block, fixed order)
__global__ void __launch_bounds__(256) k_sumd(long long n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double red[256];
  double a = 0.0;
  for (long long i = threadIdx.x; i < n; i += 256) a += x[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Adam step on the FP32 master weights, then their BF16 copy.
__global__ void __launch_bounds__(256)
    k_adam(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
           float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
           bf16_t* __restrict__ pbf) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  p[i] = v;
  pbf[i] = f2bf(v);
}

__global__ void __launch_bounds__(256) k_to_bf16(long long n, const float* __restrict__ p, bf16_t* __restrict__ q) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = f2bf(p[i]);
}

// Fourier bases of a curve's nodes: T[i][2m] = cos(m t_i), T[i][2m+1] =
// -sin(m t_i); Tw = T * wt_i / (2 pi) (wt: the parameter quadrature weight).
__global__ void __launch_bounds__(256)
    k_bases(long long N, const double* __restrict__ t, const double* __restrict__ wt, bf16_t* __restrict__ T,
            bf16_t* __restrict__ Tw) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N * kModes) return;
  const long long i = q / kModes;
  const int m = (int)(q % kModes);
  double s, c;
  sincos((double)m * t[i], &s, &c);
  const double f = wt[i] / (2.0 * M_PI);
  T[i * kS + 2 * m] = f2bf((float)c);
  T[i * kS + 2 * m + 1] = f2bf((float)-s);
  Tw[i * kS + 2 * m] = f2bf((float)(c * f));
  Tw[i * kS + 2 * m + 1] = f2bf((float)(-s * f));
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ------------------------------------------------------------ per curve ---
// What the network needs of a plan's curve, in patch order, on the device.
struct ifgf::FnoGeo {
  long long N = 0;
  int Np = 0;
  double k = 0.0;
  DevBuf<float> geo;     // 7 channels per node
  DevBuf<double> dfac;   // d/ds of a node = dfac * d/dx (x the patch's [-1, 1] variable)
  DevBuf<double> D;      // Np x Np differentiation on the Fejer nodes
  DevBuf<bf16_t> T, Tw;  // Fourier bases (N x 64)
  DevBuf<double> t, wt;
  std::vector<double> t_host, wt_host;
};
void ifgf::destroy_fno_geo(FnoGeo* g) { delete g; }

namespace {

// Chebyshev differentiation on the first-kind points x_j (Fejer nodes):
// barycentric weights w_j = (-1)^j sin((2j+1) pi / (2n)), D_jq = (w_q / w_j) /
// (x_j - x_q) off the diagonal, D_jj = -sum_{q != j} D_jq.
std::vector<double> cheb_diff(const std::vector<double>& x) {
  const int n = (int)x.size();
  std::vector<double> w(n), D((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) w[j] = ((j % 2) ? -1.0 : 1.0) * std::sin((2.0 * j + 1.0) * M_PI / (2.0 * n));
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int q = 0; q < n; q++) {
      if (q == j) continue;
      const double d = (w[q] / w[j]) / (x[j] - x[q]);
      D[(size_t)j * n + q] = d;
      s += d;
    }
    D[(size_t)j * n + j] = -s;
  }
  return D;
}

FnoGeo& fno_geo(ifgf_plan* P, hipStream_t st) {
  if (P->fno_geo) return *P->fno_geo;
  const Geometry& g = P->geo;
  const Curve& c = P->curve;
  const long long N = g.N;
  const int Np = g.Np;
  std::unique_ptr<FnoGeo> F(new FnoGeo());
  F->N = N;
  F->Np = Np;
  F->k = P->prm.k;
  double cx = 0.0, cy = 0.0, L = 0.0;
  for (long long i = 0; i < N; i++) {
    cx += g.w[i] * g.x[i];
    cy += g.w[i] * g.y[i];
    L += g.w[i];
  }
  cx /= L;
  cy /= L;
  double Rs = 0.0;
  for (long long i = 0; i < N; i++) Rs = std::max(Rs, std::hypot(g.x[i] - cx, g.y[i] - cy));
  std::vector<float> geo((size_t)N * 7);
  std::vector<double> dfac(N), tt(N), wt(N);
  double s = 0.0;
  for (long long i = 0; i < N; i++) {
    const int p = (int)(i / Np), j = (int)(i % Np);
    const double t = g.t[i], h = 1e-5;
    // curvature: d(tangent angle)/ds by central differences of x'(t)
    const double a1 = std::atan2(c.yp(t + h), c.xp(t + h)), a0 = std::atan2(c.yp(t - h), c.xp(t - h));
    double da = a1 - a0;
    if (da > M_PI) da -= 2.0 * M_PI;
    if (da < -M_PI) da += 2.0 * M_PI;
    const double speed = std::hypot(c.xp(t), c.yp(t));
    const double kappa = da / (2.0 * h) / speed;
    const double half = 0.5 * (g.t2[p] - g.t1[p]);
    wt[i] = half * g.fejer_weights[j];
    tt[i] = t;
    dfac[i] = 1.0 / (half * speed);
    const double si = s + 0.5 * g.w[i];  // arclength at the node (midpoint of its weight)
    s += g.w[i];
    float* q = &geo[(size_t)i * 7];
    q[0] = (float)((g.x[i] - cx) / Rs);
    q[1] = (float)((g.y[i] - cy) / Rs);
    q[2] = (float)g.nx[i];
    q[3] = (float)g.ny[i];
    q[4] = (float)std::max(-8.0, std::min(8.0, kappa * Rs / 4.0));
    q[5] = (float)(si / L);
    q[6] = (float)(P->prm.k * L / 2000.0);
  }
  F->geo.upload(geo);
  F->dfac.upload(dfac);
  F->D.upload(cheb_diff(g.fejer_nodes));
  F->t.upload(tt);
  F->wt.upload(wt);
  F->t_host = tt;
  F->wt_host = wt;
  F->T.alloc((size_t)N * kS);
  F->Tw.alloc((size_t)N * kS);
  hipLaunchKernelGGL(k_bases, dim3(nblk(N * kModes)), dim3(256), 0, st, N, F->t.p, F->wt.p, F->T.p, F->Tw.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  P->fno_geo = F.release();
  return *P->fno_geo;
}

// GEMM launch (see k_gemm). Split K over at most zmax blocks when the output
// is one tile and K is long (the DFT products and the weight gradients).
template <typename TA_, typename TB_, bool TRA, bool TRB>
void gemm(hipStream_t st, int M, int N, int K, const TA_* A, long long lda, const TB_* B, long long ldb, float* C,
          long long ldc, bool accumulate, const float* bias, float* part, size_t part_cap) {
  const int gm = (M + BM - 1) / BM, gn = (N + BN - 1) / BN;
  int Z = 1;
  if ((long long)gm * gn < 64 && K > 4 * BK && part) {
    Z = std::min<long long>(256, (K + 8 * BK - 1) / (8 * BK));
    while (Z > 1 && (size_t)Z * M * N > part_cap) Z--;
  }
  int kchunk = (K + Z - 1) / Z;
  kchunk = (kchunk + BK - 1) / BK * BK;
  Z = (K + kchunk - 1) / kchunk;
  if (Z <= 1) {
    hipLaunchKernelGGL((k_gemm<TA_, TB_, TRA, TRB>), dim3(gm, gn, 1), dim3(256), 0, st, M, N, K, A, lda, B, ldb, C,
                       ldc, accumulate ? 1 : 0, bias, K, (float*)nullptr);
  } else {
    hipLaunchKernelGGL((k_gemm<TA_, TB_, TRA, TRB>), dim3(gm, gn, Z), dim3(256), 0, st, M, N, K, A, lda, B, ldb, C,
                       ldc, 0, (const float*)nullptr, kchunk, part);
    hipLaunchKernelGGL(k_gemm_reduce, dim3(nblk((long long)M * N)), dim3(256), 0, st, M, N, Z, part, C, ldc,
                       accumulate ? 1 : 0, bias);
  }
}

}  // namespace

// -------------------------------------------------------------- network ---
struct FnoSample {
  long long N = 0;
  int Np = 0;
  DevBuf<float> X, mphys;
  DevBuf<zd> v, target;
  DevBuf<bf16_t> T, Tw;
  std::vector<double> t, wt;  // node parameters and parameter quadrature weights (host)
  double inv = 0.0;  // 1 / sum |target|^2
};

struct ifgf_fno {
  std::string err;
  hipStream_t st = nullptr;
  DevBuf<float> p, g, m1, m2;  // master weights, gradient, Adam moments
  DevBuf<bf16_t> pb;           // BF16 weights
  long long adam_t = 0;
  std::vector<std::unique_ptr<FnoSample>> samples;
  // workspaces (grown on demand)
  long long cap = 0;  // nodes
  int bcap = 0;       // samples per batch
  DevBuf<float> X, H[kL + 1], Pl[kL], Qp, Q, O, dO, dQ, dH, dP, mph;
  DevBuf<float> Hh, G, dG, dHh;  // [layer][B][64][64] for Hh; [B][64][64] others
  DevBuf<float> part;
  size_t part_cap = 0;
  DevBuf<double> stats, lt, lsum;
  ~ifgf_fno() {
    if (st) (void)hipStreamDestroy(st);
  }
  void reserve(long long nodes, int B) {
    if (nodes > cap) {
      cap = nodes;
      X.alloc((size_t)cap * kCinPad);
      for (auto& h : H) h.alloc((size_t)cap * kW);
      for (auto& q : Pl) q.alloc((size_t)cap * kW);
      Qp.alloc((size_t)cap * kW);
      Q.alloc((size_t)cap * kW);
      O.alloc((size_t)cap * kOutPad);
      dO.alloc((size_t)cap * kOutPad);
      dQ.alloc((size_t)cap * kW);
      dH.alloc((size_t)cap * kW);
      dP.alloc((size_t)cap * kW);
      mph.alloc(cap);
      lt.alloc(cap);
    }
    if (B > bcap) {
      bcap = B;
      Hh.alloc((size_t)kL * B * kS * kW);
      G.alloc((size_t)B * kS * kW);
      dG.alloc((size_t)B * kS * kW);
      dHh.alloc((size_t)B * kS * kW);
      lsum.alloc(B);
    }
    if (!part.p) {
      part_cap = (size_t)256 * kS * kW;
      part.alloc(part_cap);
      stats.alloc(4);
    }
  }
  const bf16_t* w(long long off) const { return pb.p + off; }
  const float* f(long long off) const { return p.p + off; }

  // Forward over a batch of nb samples whose nodes are concatenated in X
  // (segment s: nodes [off[s], off[s] + N[s]), bases T[s], Tw[s]); O gets
  // the output columns.
  void forward(int nb, const long long* off, const long long* Ns, const bf16_t* const* T, const bf16_t* const* Tw,
               long long Nt) {
    gemm<float, bf16_t, false, false>(st, (int)Nt, kW, kCinPad, X.p, kCinPad, w(kLay.wl), kW, H[0].p, kW, false,
                                      f(kLay.bl), part.p, part_cap);
    for (int l = 0; l < kL; l++) {
      float* Hhl = Hh.p + (size_t)l * bcap * kS * kW;
      for (int s = 0; s < nb; s++)
        gemm<bf16_t, float, true, false>(st, kS, kW, (int)Ns[s], Tw[s], kS, H[l].p + off[s] * kW, kW,
                                         Hhl + (size_t)s * kS * kW, kW, false, nullptr, part.p, part_cap);
      hipLaunchKernelGGL(k_mix_fwd, dim3(nblk((long long)nb * kModes * kW)), dim3(256), 0, st, nb, Hhl, w(kLay.R[l]),
                         G.p);
      for (int s = 0; s < nb; s++)
        gemm<bf16_t, float, false, false>(st, (int)Ns[s], kW, kS, T[s], kS, G.p + (size_t)s * kS * kW, kW,
                                          Pl[l].p + off[s] * kW, kW, false, nullptr, part.p, part_cap);
      gemm<float, bf16_t, false, false>(st, (int)Nt, kW, kW, H[l].p, kW, w(kLay.W[l]), kW, Pl[l].p, kW, true,
                                        f(kLay.b[l]), part.p, part_cap);
      if (l < kL - 1)
        hipLaunchKernelGGL(k_gelu_fwd, dim3(nblk(Nt * kW)), dim3(256), 0, st, Nt * kW, Pl[l].p, H[l + 1].p);
      else
        HIP_TRY(hipMemcpyAsync(H[kL].p, Pl[l].p, (size_t)Nt * kW * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    gemm<float, bf16_t, false, false>(st, (int)Nt, kW, kW, H[kL].p, kW, w(kLay.wp1), kW, Qp.p, kW, false,
                                      f(kLay.bp1), part.p, part_cap);
    hipLaunchKernelGGL(k_gelu_fwd, dim3(nblk(Nt * kW)), dim3(256), 0, st, Nt * kW, Qp.p, Q.p);
    gemm<float, bf16_t, false, false>(st, (int)Nt, kOutPad, kW, Q.p, kW, w(kLay.wp2), kOutPad, O.p, kOutPad, false,
                                      f(kLay.bp2), part.p, part_cap);
  }

  // Backward from dO (filled by the loss) into g (added to).
  void backward(int nb, const long long* off, const long long* Ns, const bf16_t* const* T, const bf16_t* const* Tw,
                long long Nt) {
    float* gg = g.p;
    // projection
    gemm<float, float, true, false>(st, kW, kOutPad, (int)Nt, Q.p, kW, dO.p, kOutPad, gg + kLay.wp2, kOutPad, true,
                                    nullptr, part.p, part_cap);
    hipLaunchKernelGGL(k_colsum, dim3(kOutPad), dim3(256), 0, st, Nt, kOutPad, dO.p, (long long)kOutPad,
                       gg + kLay.bp2, 1);
    gemm<float, bf16_t, false, true>(st, (int)Nt, kW, kOutPad, dO.p, kOutPad, w(kLay.wp2), kOutPad, dQ.p, kW, false,
                                     nullptr, part.p, part_cap);
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(Nt * kW)), dim3(256), 0, st, Nt * kW, Qp.p, dQ.p, dQ.p);
    gemm<float, float, true, false>(st, kW, kW, (int)Nt, H[kL].p, kW, dQ.p, kW, gg + kLay.wp1, kW, true, nullptr,
                                    part.p, part_cap);
    hipLaunchKernelGGL(k_colsum, dim3(kW), dim3(256), 0, st, Nt, kW, dQ.p, (long long)kW, gg + kLay.bp1, 1);
    gemm<float, bf16_t, false, true>(st, (int)Nt, kW, kW, dQ.p, kW, w(kLay.wp1), kW, dH.p, kW, false, nullptr,
                                     part.p, part_cap);
    // Fourier layers, last to first; dH holds dL/dH_{l+1}
    for (int l = kL - 1; l >= 0; l--) {
      if (l < kL - 1)
        hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(Nt * kW)), dim3(256), 0, st, Nt * kW, Pl[l].p, dH.p, dP.p);
      else
        HIP_TRY(hipMemcpyAsync(dP.p, dH.p, (size_t)Nt * kW * sizeof(float), hipMemcpyDeviceToDevice, st));
      gemm<float, float, true, false>(st, kW, kW, (int)Nt, H[l].p, kW, dP.p, kW, gg + kLay.W[l], kW, true, nullptr,
                                      part.p, part_cap);
      hipLaunchKernelGGL(k_colsum, dim3(kW), dim3(256), 0, st, Nt, kW, dP.p, (long long)kW, gg + kLay.b[l], 1);
      gemm<float, bf16_t, false, true>(st, (int)Nt, kW, kW, dP.p, kW, w(kLay.W[l]), kW, dH.p, kW, false, nullptr,
                                       part.p, part_cap);
      // spectral path: dG = T^T dP per sample, then the mixing, then dH += Tw dHhat
      for (int s = 0; s < nb; s++)
        gemm<bf16_t, float, true, false>(st, kS, kW, (int)Ns[s], T[s], kS, dP.p + off[s] * kW, kW,
                                         dG.p + (size_t)s * kS * kW, kW, false, nullptr, part.p, part_cap);
      const float* Hhl = Hh.p + (size_t)l * bcap * kS * kW;
      hipLaunchKernelGGL(k_mix_bwd_w, dim3(nblk((long long)kModes * kW * kW)), dim3(256), 0, st, nb, Hhl, dG.p,
                         gg + kLay.R[l]);
      hipLaunchKernelGGL(k_mix_bwd_in, dim3(nblk((long long)nb * kModes * kW)), dim3(256), 0, st, nb, dG.p,
                         w(kLay.R[l]), dHh.p);
      for (int s = 0; s < nb; s++)
        gemm<bf16_t, float, false, false>(st, (int)Ns[s], kW, kS, Tw[s], kS, dHh.p + (size_t)s * kS * kW, kW,
                                          dH.p + off[s] * kW, kW, true, nullptr, part.p, part_cap);
    }
    // lifting (no input gradient needed)
    gemm<float, float, true, false>(st, kCinPad, kW, (int)Nt, X.p, kCinPad, dH.p, kW, gg + kLay.wl, kW, true,
                                    nullptr, part.p, part_cap);
    hipLaunchKernelGGL(k_colsum, dim3(kW), dim3(256), 0, st, Nt, kW, dH.p, (long long)kW, gg + kLay.bl, 1);
  }

  // Loss of samples ids[0..nb) with the current weights; with grad, the
  // gradient of the mean loss goes into g (cleared first). Returns the mean.
  double loss_grad(const int* ids, int nb, bool grad) {
    std::vector<long long> off(nb), Ns(nb);
    std::vector<const bf16_t*> T(nb), Tw(nb);
    long long Nt = 0;
    for (int s = 0; s < nb; s++) {
      const FnoSample& S = *samples[ids[s]];
      off[s] = Nt;
      Ns[s] = S.N;
      T[s] = S.T.p;
      Tw[s] = S.Tw.p;
      Nt += S.N;
    }
    reserve(Nt, nb);
    for (int s = 0; s < nb; s++) {
      const FnoSample& S = *samples[ids[s]];
      HIP_TRY(hipMemcpyAsync(X.p + off[s] * kCinPad, S.X.p, (size_t)S.N * kCinPad * sizeof(float),
                             hipMemcpyDeviceToDevice, st));
    }
    forward(nb, off.data(), Ns.data(), T.data(), Tw.data(), Nt);
    for (int s = 0; s < nb; s++) {
      const FnoSample& S = *samples[ids[s]];
      hipLaunchKernelGGL(k_loss, dim3(nblk(S.N)), dim3(256), 0, st, S.N, O.p + off[s] * kOutPad, S.mphys.p, S.v.p,
                         S.target.p, S.inv, 1.0 / nb, dO.p + off[s] * kOutPad, lt.p + off[s]);
      hipLaunchKernelGGL(k_sumd, dim3(1), dim3(256), 0, st, S.N, lt.p + off[s], lsum.p + s);
    }
    if (grad) {
      HIP_TRY(hipMemsetAsync(g.p, 0, (size_t)kLay.total * sizeof(float), st));
      backward(nb, off.data(), Ns.data(), T.data(), Tw.data(), Nt);
    }
    std::vector<double> ls(nb);
    HIP_TRY(hipMemcpyAsync(ls.data(), lsum.p, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double sum = 0.0;
    for (double v : ls) sum += v;
    return sum / nb;
  }

  // z = M(v) for a plan's curve (device pointers, stream-ordered on s)
  void apply(ifgf_plan* P, const zd* v, zd* z, hipStream_t s) {
    FnoGeo& F = fno_geo(P, st);
    const long long N = F.N;
    reserve(N, 1);
    hipStream_t keep = st;
    st = s;
    hipLaunchKernelGGL(k_sumsq, dim3(1), dim3(256), 0, st, N, v, stats.p);
    hipLaunchKernelGGL(k_features, dim3(nblk(N)), dim3(256), 0, st, N, F.Np, F.geo.p, F.dfac.p, F.D.p, v, stats.p,
                       F.k, X.p, mph.p);
    hipLaunchKernelGGL(k_direction_sums, dim3(1), dim3(256), 0, st, N, F.geo.p, X.p, stats.p);
    hipLaunchKernelGGL(k_direction_feature, dim3(nblk(N)), dim3(256), 0, st, N, F.geo.p, stats.p, X.p);
    const long long off = 0;
    const bf16_t* T = F.T.p;
    const bf16_t* Tw = F.Tw.p;
    forward(1, &off, &N, &T, &Tw, N);
    hipLaunchKernelGGL(k_output, dim3(nblk(N)), dim3(256), 0, st, N, O.p, mph.p, v, z);
    st = keep;
  }
};

namespace ifgf {
int fno_apply_device(ifgf_fno* F, ifgf_plan* P, const zd* v, zd* z, hipStream_t st) {
  try {
    F->apply(P, v, z, st);
    HIP_TRY(hipGetLastError());
  } catch (std::exception& e) {
    return plan_fail(P, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}
}  // namespace ifgf

// ------------------------------------------------------------------ C ABI ---
namespace {
thread_local std::string g_fno_err;
int fno_fail(ifgf_fno* F, int code, const std::string& msg) {
  g_fno_err = msg;
  if (F) F->err = msg;
  return code;
}

// splitmix64 (Vigna) -> uniform in [-1, 1)
inline double sm_uniform(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
}  // namespace

extern "C" {

long long ifgf_fno_num_params(void) { return kLay.total; }

const char* ifgf_fno_error(const ifgf_fno* F) { return F ? F->err.c_str() : g_fno_err.c_str(); }

int ifgf_fno_create(unsigned long long seed, ifgf_fno** out) {
  if (!out) return fno_fail(nullptr, IFGF_ERR_ARG, "null argument");
  *out = nullptr;
  try {
    std::unique_ptr<ifgf_fno> F(new ifgf_fno());
    HIP_TRY(hipStreamCreate(&F->st));
    // uniform initialisation: Glorot bounds for the dense weights, 1 / (in
    // out) for the spectral weights (the usual FNO scale), zero biases; the
    // last projection starts small so the untrained multiplier is m_phys
    std::vector<float> h(kLay.total, 0.f);
    uint64_t idx = 0;
    auto fill = [&](long long o, long long n, double a) {
      for (long long i = 0; i < n; i++) h[o + i] = (float)(a * sm_uniform(seed, idx++));
    };
    fill(kLay.wl, (long long)kCinPad * kW, std::sqrt(6.0 / (kCin + kW)));
    for (int l = 0; l < kL; l++) {
      fill(kLay.R[l], (long long)kModes * kW * kW * 2, 1.0 / (kW * kW) * 8.0);
      fill(kLay.W[l], kW * kW, std::sqrt(6.0 / (2 * kW)));
    }
    fill(kLay.wp1, kW * kW, std::sqrt(6.0 / (2 * kW)));
    fill(kLay.wp2, kW * kOutPad, 0.01 * std::sqrt(6.0 / (kW + 2)));
    // rows of Wlift for the padded input channels stay zero
    for (int r = kCin; r < kCinPad; r++)
      for (int c = 0; c < kW; c++) h[kLay.wl + r * kW + c] = 0.f;
    F->p.upload(h);
    F->pb.alloc(kLay.total);
    F->g.alloc(kLay.total);
    F->m1.upload(std::vector<float>(kLay.total, 0.f));
    F->m2.upload(std::vector<float>(kLay.total, 0.f));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(kLay.total)), dim3(256), 0, F->st, kLay.total, F->p.p, F->pb.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(F->st));
    *out = F.release();
  } catch (std::exception& e) {
    return fno_fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

void ifgf_fno_destroy(ifgf_fno* F) { delete F; }

int ifgf_fno_get_params(const ifgf_fno* F, float* out) {
  if (!F || !out) return fno_fail(nullptr, IFGF_ERR_ARG, "null argument");
  try {
    HIP_TRY(hipMemcpy(out, F->p.p, kLay.total * sizeof(float), hipMemcpyDeviceToHost));
  } catch (std::exception& e) {
    return fno_fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_fno_set_params(ifgf_fno* F, const float* in) {
  if (!F || !in) return fno_fail(F, IFGF_ERR_ARG, "null argument");
  try {
    HIP_TRY(hipMemcpy(F->p.p, in, kLay.total * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(kLay.total)), dim3(256), 0, F->st, kLay.total, F->p.p, F->pb.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(F->st));
  } catch (std::exception& e) {
    return fno_fail(F, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_fno_add_sample(ifgf_fno* F, ifgf_plan* P, const double* v, const double* target) {
  if (!F || !P || !v || !target) return fno_fail(F, IFGF_ERR_ARG, "null argument");
  if (!P->on_device || P->field) return fno_fail(F, IFGF_ERR_NO_DEVICE, "plan was built host-only or is a field");
  try {
    FnoGeo& G = fno_geo(P, F->st);
    std::unique_ptr<FnoSample> S(new FnoSample());
    const long long N = G.N;
    S->N = N;
    S->Np = G.Np;
    S->v.upload(std::vector<zd>((const zd*)v, (const zd*)v + N));
    S->target.upload(std::vector<zd>((const zd*)target, (const zd*)target + N));
    double nt = 0.0;
    for (long long i = 0; i < 2 * N; i++) nt += target[i] * target[i];
    if (!(nt > 0.0)) return fno_fail(F, IFGF_ERR_ARG, "zero target density");
    S->inv = 1.0 / nt;
    S->X.alloc((size_t)N * kCinPad);
    S->mphys.alloc(N);
    S->T.alloc((size_t)N * kS);
    S->Tw.alloc((size_t)N * kS);
    F->reserve(N, 1);
    hipLaunchKernelGGL(k_sumsq, dim3(1), dim3(256), 0, F->st, N, S->v.p, F->stats.p);
    hipLaunchKernelGGL(k_features, dim3(nblk(N)), dim3(256), 0, F->st, N, G.Np, G.geo.p, G.dfac.p, G.D.p, S->v.p,
                       F->stats.p, G.k, S->X.p, S->mphys.p);
    hipLaunchKernelGGL(k_direction_sums, dim3(1), dim3(256), 0, F->st, N, G.geo.p, S->X.p, F->stats.p);
    hipLaunchKernelGGL(k_direction_feature, dim3(nblk(N)), dim3(256), 0, F->st, N, G.geo.p, F->stats.p, S->X.p);
    HIP_TRY(hipMemcpyAsync(S->T.p, G.T.p, (size_t)N * kS * sizeof(bf16_t), hipMemcpyDeviceToDevice, F->st));
    HIP_TRY(hipMemcpyAsync(S->Tw.p, G.Tw.p, (size_t)N * kS * sizeof(bf16_t), hipMemcpyDeviceToDevice, F->st));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(F->st));
    S->t = G.t_host;
    S->wt = G.wt_host;
    F->samples.push_back(std::move(S));
  } catch (std::exception& e) {
    return fno_fail(F, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_fno_num_samples(const ifgf_fno* F) { return F ? (int)F->samples.size() : -1; }

int ifgf_fno_clear_samples(ifgf_fno* F) {
  if (!F) return fno_fail(nullptr, IFGF_ERR_ARG, "null argument");
  F->samples.clear();
  return IFGF_OK;
}

// Features, multiplier prior, node parameters and parameter quadrature
// weights of sample s (host copies; any output may be NULL): X [N][11],
// mphys [N], t [N], wt [N].
int ifgf_fno_sample_data(const ifgf_fno* F, int s, long long* N, float* X, float* mphys, double* t, double* wt) {
  if (!F || s < 0 || s >= (int)F->samples.size() || !N) return fno_fail(nullptr, IFGF_ERR_ARG, "invalid argument");
  const FnoSample& S = *F->samples[s];
  *N = S.N;
  try {
    if (X) {
      std::vector<float> h((size_t)S.N * kCinPad);
      HIP_TRY(hipMemcpy(h.data(), S.X.p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (long long i = 0; i < S.N; i++)
        for (int c = 0; c < kCin; c++) X[i * kCin + c] = h[(size_t)i * kCinPad + c];
    }
    if (mphys) HIP_TRY(hipMemcpy(mphys, S.mphys.p, S.N * sizeof(float), hipMemcpyDeviceToHost));
    if (t) std::copy(S.t.begin(), S.t.end(), t);
    if (wt) std::copy(S.wt.begin(), S.wt.end(), wt);
  } catch (std::exception& e) {
    return fno_fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_fno_loss_grad(ifgf_fno* F, const int* ids, int n, double* loss, float* grad) {
  if (!F || !ids || n < 1 || !loss) return fno_fail(F, IFGF_ERR_ARG, "invalid argument");
  for (int s = 0; s < n; s++)
    if (ids[s] < 0 || ids[s] >= (int)F->samples.size()) return fno_fail(F, IFGF_ERR_ARG, "sample index out of range");
  try {
    *loss = F->loss_grad(ids, n, grad != nullptr);
    if (grad) HIP_TRY(hipMemcpy(grad, F->g.p, kLay.total * sizeof(float), hipMemcpyDeviceToHost));
  } catch (std::exception& e) {
    return fno_fail(F, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_fno_train(ifgf_fno* F, const int* ids, int n, int epochs, int batch, double lr, unsigned long long seed,
                   double* loss_hist) {
  if (!F || !ids || n < 1 || epochs < 1 || batch < 1 || !(lr > 0.0)) return fno_fail(F, IFGF_ERR_ARG, "invalid argument");
  for (int s = 0; s < n; s++)
    if (ids[s] < 0 || ids[s] >= (int)F->samples.size()) return fno_fail(F, IFGF_ERR_ARG, "sample index out of range");
  try {
    std::vector<int> order(ids, ids + n);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    uint64_t draw = 0;
    for (int ep = 0; ep < epochs; ep++) {
      // seeded shuffle (Fisher-Yates on splitmix64 draws)
      for (int i = n - 1; i > 0; i--) {
        const double u = 0.5 * (sm_uniform(seed, draw++) + 1.0);
        const int j = std::min(i, (int)(u * (i + 1)));
        std::swap(order[i], order[j]);
      }
      double esum = 0.0;
      int nbt = 0;
      for (int b0 = 0; b0 < n; b0 += batch) {
        const int nb = std::min(batch, n - b0);
        esum += F->loss_grad(order.data() + b0, nb, true);
        nbt++;
        F->adam_t++;
        const float c1 = 1.f - std::pow(b1, (float)F->adam_t), c2 = 1.f - std::pow(b2, (float)F->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nblk(kLay.total)), dim3(256), 0, F->st, kLay.total, F->p.p, F->g.p, F->m1.p,
                           F->m2.p, (float)lr, b1, b2, eps, c1, c2, F->pb.p);
      }
      if (loss_hist) loss_hist[ep] = esum / nbt;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(F->st));
  } catch (std::exception& e) {
    return fno_fail(F, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_fno_apply(ifgf_fno* F, ifgf_plan* P, const double* v, double* z, int flags, void* stream) {
  if (!F || !P || !v || !z) return fno_fail(F, IFGF_ERR_ARG, "null argument");
  if (!P->on_device || P->field) return fno_fail(F, IFGF_ERR_NO_DEVICE, "plan was built host-only or is a field");
  try {
    const long long N = P->geo.N;
    if (flags & IFGF_PTR_DEVICE) {
      F->apply(P, (const zd*)v, (zd*)z, stream ? (hipStream_t)stream : F->st);
      HIP_TRY(hipGetLastError());
    } else {
      DevBuf<zd> dv, dz;
      dv.upload(std::vector<zd>((const zd*)v, (const zd*)v + N));
      dz.alloc(N);
      F->apply(P, dv.p, dz.p, F->st);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(z, dz.p, N * sizeof(zd), hipMemcpyDeviceToHost, F->st));
      HIP_TRY(hipStreamSynchronize(F->st));
    }
  } catch (std::exception& e) {
    return fno_fail(F, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

}  // extern "C"

// BF16 GEMM on the MFMA for tests of the kernel itself: C (M x N) = A (M x K)
// B (K x N), FP32 row-major host arrays (rounded to BF16 on the device).
extern "C" int ifgf_fno_gemm_check(int M, int N, int K, int trans_a, int trans_b, const float* A, const float* B,
                                   float* C) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || !C) return fno_fail(nullptr, IFGF_ERR_ARG, "invalid argument");
  try {
    DevBuf<float> a, b, c, part;
    a.upload(std::vector<float>(A, A + (size_t)M * K));
    b.upload(std::vector<float>(B, B + (size_t)K * N));
    c.alloc((size_t)M * N);
    const size_t cap = (size_t)256 * M * N;
    part.alloc(std::min<size_t>(cap, (size_t)1 << 26));
    const long long lda = trans_a ? M : K, ldb = trans_b ? K : N;
    if (!trans_a && !trans_b)
      gemm<float, float, false, false>(0, M, N, K, a.p, lda, b.p, ldb, c.p, N, false, nullptr, part.p, part.n);
    else if (trans_a && !trans_b)
      gemm<float, float, true, false>(0, M, N, K, a.p, lda, b.p, ldb, c.p, N, false, nullptr, part.p, part.n);
    else if (!trans_a && trans_b)
      gemm<float, float, false, true>(0, M, N, K, a.p, lda, b.p, ldb, c.p, N, false, nullptr, part.p, part.n);
    else
      return fno_fail(nullptr, IFGF_ERR_ARG, "both operands transposed is not instantiated");
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(C, c.p, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost));
  } catch (std::exception& e) {
    return fno_fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}
