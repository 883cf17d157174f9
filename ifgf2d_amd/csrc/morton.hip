// Device Morton box assignment and stable radix sort of the points
// (BoxTree::SortBox, IFGF_Source/BoxTree.h:73-106, with the bounding box of
// BoxTree::ComputeBB, BoxTree.h:179-202, and Level::Point2Morton,
// Level.h:145-161).
//
//   k_bbox_partial / k_bbox_final   exact min / max of the coordinates, then the
//                                   reference's +-1e-10 margins and the leaf box size
//   k_morton_codes                  leaf box (floor((x - min) / h_leaf), same FP64
//                                   operations as the host) and interleaved code
//   k_radix_hist / k_radix_scan / k_radix_scatter
//                                   LSD radix sort, 8 bits per pass, stable: a key's
//                                   rank among equal digits follows its input order,
//                                   so ties keep input order like std::stable_sort
//
// The result is bit-identical to the host's stable sort by construction (the
// same IEEE operations on the same inputs, an order-preserving sort); the
// tests compare the two.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/ifgf2d.h"
#include "plan_internal.h"

using namespace ifgf;

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;                    // keys per thread per tile
constexpr int kTile = kThreads * kItems;     // keys per block
constexpr int kRadix = 256;
constexpr int kWaves = kThreads / 64;

__global__ void __launch_bounds__(kThreads)
    k_bbox_partial(long long n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ part) {
  double mnx = 1e308, mny = 1e308, mxx = -1e308, mxy = -1e308;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    mnx = fmin(mnx, x[i]);
    mny = fmin(mny, y[i]);
    mxx = fmax(mxx, x[i]);
    mxy = fmax(mxy, y[i]);
  }
  __shared__ double s[4][kThreads];
  s[0][threadIdx.x] = mnx;
  s[1][threadIdx.x] = mny;
  s[2][threadIdx.x] = mxx;
  s[3][threadIdx.x] = mxy;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + o]);
      s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + o]);
      s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + o]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[(size_t)blockIdx.x * 4 + threadIdx.x] = s[threadIdx.x][0];
}

// box[0..3] = min x - 1e-10, min y - 1e-10, leaf box size, (unused)
__global__ void __launch_bounds__(64) k_bbox_final(int nblk, const double* __restrict__ part, int D,
                                                   double* __restrict__ box) {
  if (threadIdx.x != 0) return;
  double mn0 = 1e308, mn1 = 1e308, mx0 = -1e308, mx1 = -1e308;
  for (int b = 0; b < nblk; b++) {
    mn0 = fmin(mn0, part[b * 4 + 0]);
    mn1 = fmin(mn1, part[b * 4 + 1]);
    mx0 = fmax(mx0, part[b * 4 + 2]);
    mx1 = fmax(mx1, part[b * 4 + 3]);
  }
  mx0 += 1e-10;
  mx1 += 1e-10;
  mn0 -= 1e-10;
  mn1 -= 1e-10;
  const double bs0 = fmax(mx0 - mn0, mx1 - mn1);
  box[0] = mn0;
  box[1] = mn1;
  box[2] = bs0 / (double)(1ll << D);
  box[3] = bs0;
}

__device__ inline uint32_t spread_bits(uint32_t v) {  // Level.h:172-237 (splitBy2)
  uint64_t x = v;
  x = (x | (x << 16)) & 0xFFFF0000FFFFull;
  x = (x | (x << 8)) & 0xFF00FF00FF00FFull;
  x = (x | (x << 4)) & 0xF0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return (uint32_t)x;
}

// code[i] = interleave(ix, iy), x in the low bit; bad[0] = 1 if a point falls
// outside the 2^D x 2^D leaf grid (Level.h:153-157).
__global__ void __launch_bounds__(kThreads)
    k_morton_codes(long long n, const double* __restrict__ x, const double* __restrict__ y,
                   const double* __restrict__ box, int D, uint32_t* __restrict__ code, int* __restrict__ idx,
                   int* __restrict__ bad) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const double mn0 = box[0], mn1 = box[1], bs = box[2];
  const long long px = (long long)((x[i] - mn0) / bs), py = (long long)((y[i] - mn1) / bs);
  const long long nb = 1ll << D;
  if (px < 0 || py < 0 || px >= nb || py >= nb) {
    bad[0] = 1;
    code[i] = 0xFFFFFFFFu;
  } else {
    code[i] = spread_bits((uint32_t)px) | (spread_bits((uint32_t)py) << 1);
  }
  idx[i] = (int)i;
}

// hist[d * nblk + b] = keys of tile b whose digit (key >> shift) & 255 is d
__global__ void __launch_bounds__(kThreads)
    k_radix_hist(long long n, const uint32_t* __restrict__ key, int shift, int nblk, unsigned* __restrict__ hist) {
  __shared__ unsigned cnt[kRadix];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const long long t0 = (long long)blockIdx.x * kTile;
  for (int r = 0; r < kItems; r++) {
    const long long i = t0 + (long long)r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(key[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// exclusive prefix sum of hist[0..m) in place, one block of 1024 threads
__global__ void __launch_bounds__(1024) k_radix_scan(long long m, unsigned* __restrict__ hist) {
  __shared__ unsigned tot[1024];
  const long long per = (m + 1023) / 1024;
  const long long a = threadIdx.x * per, b = a + per < m ? a + per : m;
  unsigned s = 0;
  for (long long i = a; i < b; i++) s += hist[i];
  tot[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan of the chunk sums
    const unsigned v = (int)threadIdx.x >= o ? tot[threadIdx.x - o] : 0u;
    __syncthreads();
    tot[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned run = threadIdx.x ? tot[threadIdx.x - 1] : 0u;
  for (long long i = a; i < b; i++) {
    const unsigned v = hist[i];
    hist[i] = run;
    run += v;
  }
}

// Stable scatter of tile b: rounds of kThreads consecutive keys; within a round
// a key's rank among equal digits = earlier rounds + lower waves + lower lanes.
__global__ void __launch_bounds__(kThreads)
    k_radix_scatter(long long n, const uint32_t* __restrict__ kin, const int* __restrict__ vin, int shift, int nblk,
                    const unsigned* __restrict__ offs, uint32_t* __restrict__ kout, int* __restrict__ vout) {
  __shared__ unsigned base[kRadix];
  __shared__ unsigned wcnt[kWaves][kRadix];
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  base[tid] = offs[(size_t)tid * nblk + blockIdx.x];
  for (int w = 0; w < kWaves; w++) wcnt[w][tid] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const long long t0 = (long long)blockIdx.x * kTile;
  for (int r = 0; r < kItems; r++) {
    const long long i = t0 + (long long)r * kThreads + tid;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const int v = valid ? vin[i] : 0;
    const unsigned d = (k >> shift) & 255u;
    // lanes of this wave with the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const unsigned rank = (unsigned)__popcll(peers & lt);
    // the lowest lane of each digit group publishes the group's count
    if (valid && (peers & lt) == 0ull) wcnt[wave][d] = (unsigned)__popcll(peers);
    __syncthreads();
    if (valid) {
      unsigned pos = base[d] + rank;
      for (int w = 0; w < wave; w++) pos += wcnt[w][d];
      if (pos < n) {  // always true for a consistent histogram; keeps the store in bounds regardless
        kout[pos] = k;
        vout[pos] = v;
      }
    }
    __syncthreads();
    unsigned add = 0;
    for (int w = 0; w < kWaves; w++) {
      add += wcnt[w][tid];
      wcnt[w][tid] = 0;
    }
    base[tid] += add;
    __syncthreads();
  }
}

}  // namespace

// Host driver (device buffers; synchronous on st). perm[i] = input index of
// the i-th point in Morton order; codes (may be null) the leaf Morton codes
// in input order. Returns false when a point lies outside the leaf grid.
bool ifgf::device_morton_sort(const double* x, const double* y, long long n, int nlevels, std::vector<int>& perm,
                              std::vector<long long>* codes, hipStream_t st) {
  const int D = nlevels - 1;
  DevBuf<double> dx, dy, part, box;
  DevBuf<uint32_t> k0, k1;
  DevBuf<int> v0, v1, bad;
  DevBuf<unsigned> hist;
  dx.upload(std::vector<double>(x, x + n));
  dy.upload(std::vector<double>(y, y + n));
  const int pblk = (int)std::min<long long>(256, (n + kThreads - 1) / kThreads);
  part.alloc((size_t)pblk * 4);
  box.alloc(4);
  k0.alloc(n);
  k1.alloc(n);
  v0.alloc(n);
  v1.alloc(n);
  bad.alloc(1);
  const int nblk = (int)((n + kTile - 1) / kTile);
  hist.alloc((size_t)kRadix * nblk);
  HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_bbox_partial, dim3(pblk), dim3(kThreads), 0, st, n, dx.p, dy.p, part.p);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, st, pblk, part.p, D, box.p);
  hipLaunchKernelGGL(k_morton_codes, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, n, dx.p,
                     dy.p, box.p, D, k0.p, v0.p, bad.p);
  std::vector<uint32_t> c32;
  if (codes) {
    c32.resize(n);
    HIP_TRY(hipMemcpyAsync(c32.data(), k0.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  }
  const int bits = 2 * D;
  uint32_t* kin = k0.p;
  uint32_t* kout = k1.p;
  int* vin = v0.p;
  int* vout = v1.p;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(kThreads), 0, st, n, kin, shift, nblk, hist.p);
    hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, st, (long long)kRadix * nblk, hist.p);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(kThreads), 0, st, n, kin, vin, shift, nblk, hist.p, kout,
                       vout);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  HIP_TRY(hipGetLastError());
  int hbad = 0;
  perm.resize(n);
  HIP_TRY(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(perm.data(), vin, n * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (codes) codes->assign(c32.begin(), c32.end());
  return hbad == 0;
}

extern "C" int ifgf_morton_sort(const double* x, const double* y, long long n, int nlevels, long long* sorting,
                                long long* codes) {
  if (!x || !y || !sorting || n < 1 || n > (1ll << 30)) return plan_fail(nullptr, IFGF_ERR_ARG, "invalid argument");
  if (nlevels < 3) return plan_fail(nullptr, IFGF_ERR_LEVELS, "IFGF accelerator requires at least 3 levels.");
  if (nlevels > 16) return plan_fail(nullptr, IFGF_ERR_ARG, "nlevels above 16 (leaf codes are 32-bit)");
  try {
    hipStream_t st = nullptr;
    std::vector<int> perm;
    std::vector<long long> c;
    const bool ok = device_morton_sort(x, y, n, nlevels, perm, codes ? &c : nullptr, st);
    if (!ok) return plan_fail(nullptr, IFGF_ERR_GEOMETRY, "The 2D box index cannot be < 0 or larger than the number of boxes");
    for (long long i = 0; i < n; i++) sorting[i] = perm[i];
    if (codes)
      for (long long i = 0; i < n; i++) codes[i] = c[i];
  } catch (std::exception& e) {
    return plan_fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}
