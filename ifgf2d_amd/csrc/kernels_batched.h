// Several right-hand sides per application (ifgf_apply with nrhs > 1 on a
// stored-operator plan): the same stored near-field values, leaf operator and
// push geometry as the single-column kernels in kernels.h, each loaded once
// and applied to R columns. Densities and every intermediate are laid out
// point-major with the R columns innermost (rho_b[i * R + r], partial sums
// [item][..][r], slots [slot][r]), so a column block of one point is one
// contiguous R x 16-byte (FP64) or R x 8-byte (FP32) segment.
//
// Each kernel performs, per column, the operations of its single-column
// counterpart in the same order (the same summation order everywhere), so a
// column of a batched application agrees with a single application of that
// column to rounding (the compiler may fuse multiply-adds differently in the
// two kernels), and batched applications are bitwise deterministic.
#pragma once

#include <type_traits>

#include "kernels.h"

namespace ifgf {

// Cones per batched push task: two LDS windows of CB x P x R complex values
// within 40 KB, and at most 256 threads in the cone phase.
// entries per batched push task with unit evaluation (cones of <= 35 nodes)
// (FP64: 512, so that the staged records, 32 B each, fit the transform's
// LDS window and three blocks fit a CU; FP32: 1,024)
constexpr int batched_entries(int P, int zbytes) { return P <= 35 ? (zbytes == 16 ? 512 : 1024) : 256; }

constexpr int batched_cones(int P, int R, int zbytes) {
  const int c = 40960 / (2 * P * R * zbytes);
  const int m = c < 256 / P ? c : 256 / P;
  return m < 1 ? 1 : m;
}

// rho_b[i * R + r] = x[r * ld + perm[i]] * w[perm[i]], and the columns
// transposed in patch order for the finalize's correction: xt[o * R + r] =
// x[r * ld + o], o = perm[i]
template <int R>
__global__ void __launch_bounds__(256) k_gather_b(int N, const zd* __restrict__ x, long long ld,
                                                  const double* __restrict__ w, const int* __restrict__ perm,
                                                  zd* __restrict__ rho, zd* __restrict__ xt) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * R) return;
  const int i = (int)(t / R), r = (int)(t % R);
  const int o = perm[i];
  const zd v = x[(size_t)r * ld + o];
  rho[t] = cscale(v, w[o]);
  xt[(size_t)o * R + r] = v;
}

// k_near_items with R columns: lane (t, sl) keeps R accumulators.
template <int R>
__global__ void __launch_bounds__(256)
    k_near_items_b(int N, int nitems, const NearItem* __restrict__ items, const int4* __restrict__ nf_seg,
                   const zd* __restrict__ nf_val, const zd* __restrict__ rho, zd* __restrict__ npart) {
  const int w = xcd_block() * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  const int t = lane / kNearChunk, sl = lane % kNearChunk;
  zd acc[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = zd{0.0, 0.0};
  if (w < nitems) {
    const NearItem it = items[w];
    const int i = it.g * kNearGroup + t;
    if (i < N) {
      const dbl2* __restrict__ v = reinterpret_cast<const dbl2*>(nf_val) + it.off + t * kNearChunk + sl;
      dbl2 val[kNearItemChunks];
#pragma unroll
      for (int u = 0; u < kNearItemChunks; u++)
        if (u < it.nc) val[u] = v[u * kNearGroup * kNearChunk];
      const int4* sg = nf_seg + (size_t)i * (kNbSegInts / 4);
      const int4 q0 = sg[0], q1 = sg[1], q2 = sg[2], q3 = sg[3], q4 = sg[4];
      const int end[kMaxNb - 1] = {q0.x, q0.z, q1.x, q1.z, q2.x, q2.z, q3.x, q3.z};
      const int sh[kMaxNb] = {q0.y, q0.w, q1.y, q1.w, q2.y, q2.w, q3.y, q3.w, q4.y};
      const int len = q4.x;
#pragma unroll
      for (int u = 0; u < kNearItemChunks; u++) {
        const int o = (it.u0 + u) * kNearChunk + sl;
        if (u < it.nc && o < len) {
          int src = o + sh[0];
#pragma unroll
          for (int q = 1; q < kMaxNb; q++) src = (o >= end[q - 1]) ? o + sh[q] : src;
          const dbl2* __restrict__ rr = reinterpret_cast<const dbl2*>(rho) + (size_t)src * R;
#pragma unroll
          for (int r = 0; r < R; r++) {
            const dbl2 x = rr[r];
            cfma(acc[r], zd{val[u].x, val[u].y}, zd{x.x, x.y});
          }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = lane_group_sum<kNearChunk>(acc[r]);
  if (w < nitems && sl == 0) {
#pragma unroll
    for (int r = 0; r < R; r++) npart[((size_t)w * kNearGroup + t) * R + r] = acc[r];
  }
}

// k_leaf_items with R columns.
template <int P, int R, typename Z>
__global__ void __launch_bounds__(256)
    k_leaf_items_b(int nitems, const LeafItem* __restrict__ items, const Z* __restrict__ lf, const zd* __restrict__ rho,
                   Z* __restrict__ lpart) {
  constexpr int S = kLeafChunk;
  using V2 = typename vec2_of<Z>::type;
  const long long gt = (long long)xcd_block() * blockDim.x + threadIdx.x;
  const long long item = gt / (P * S);
  const int rm = (int)(gt % (P * S));
  const int m = rm / S, sl = rm % S;
  Z acc[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = Z{0, 0};
  if (item < nitems) {
    const LeafItem it = items[item];
    const V2* __restrict__ L = reinterpret_cast<const V2*>(lf) + it.off + m * S + sl;
    V2 val[kLeafItemTiles];
#pragma unroll
    for (int u = 0; u < kLeafItemTiles; u++)
      if (u * S + sl < it.n) val[u] = L[u * S * P];
#pragma unroll
    for (int u = 0; u < kLeafItemTiles; u++) {
      if (u * S + sl < it.n) {
        const zd* __restrict__ rr = rho + (size_t)(it.s0 + u * S + sl) * R;
#pragma unroll
        for (int r = 0; r < R; r++) zfma(acc[r], Z{val[u].x, val[u].y}, zcast<Z>(rr[r]));
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    acc[r].x = lane_sum<S>(acc[r].x);
    acc[r].y = lane_sum<S>(acc[r].y);
  }
  if (item < nitems && sl == 0) {
#pragma unroll
    for (int r = 0; r < R; r++) lpart[(item * P + m) * R + r] = acc[r];
  }
}

// k_level_push with R columns. Tasks hold at most CB cones (their R-column
// coefficients fit the LDS windows) and batched_entries(P, sizeof(Z)) entries;
// the cone phase runs on threads (cone, node) < CB * P. The evaluation runs
// on units: every cone's run of entries is cut into pieces of at most K
// entries (K chosen per task so that all pieces fit 256 / R units), and the
// R threads of a unit, one per column, hold their column's P coefficients
// in registers and walk the piece's entries, whose records the block keeps
// in LDS: one LDS read of each coefficient per piece instead of one per
// entry (cones of up to 35 nodes; larger ones evaluate one entry per thread
// from the LDS coefficients). Every slot of pbuf holds a.rs >= R columns whatever the block
// width, so the zero slot that absent children read stays zero for every
// width.
template <int Ps, int Pang, int R, int CB, typename RT>
__global__ void __launch_bounds__(256) k_level_push_b(PushArgsT<RT> a) {
  using Z = typename complex_of<RT>::type;
  constexpr int P = Ps * Pang;
  static_assert(CB * P <= 256, "cone phase needs CB * P threads");
  // unit evaluation while a column's P coefficients fit the registers
  constexpr bool kUnitEval = P <= 35;
  constexpr int kUnits = 256 / R;
  static_assert(!kUnitEval || kUnits > CB, "every cone needs a unit");
  // tmp is free after the transform and then holds the task's records
  constexpr int kRec = batched_entries(P, (int)sizeof(Z)) / 256;  // records per thread
  constexpr int kGeoZ = (int)(256 * kRec * sizeof(Geo4T<RT>) / sizeof(Z));
  constexpr int kTmp = CB * P * R > kGeoZ ? CB * P * R : kGeoZ;
  __shared__ ChebTablesT<RT> T;
  __shared__ Z buf[CB * P * R];
  __shared__ Z tmp[kTmp];
  const PushTask task = a.tasks[xcd_block()];
  const int tid = threadIdx.x;
  if (!a.leaf) load_transform_rows<Ps, Pang>(T, a.tabs);
  const int cl = tid / P, j = tid % P;
  const bool act = cl < task.nc && cl < CB;
  const long long g = (long long)task.c0 + cl;
  Geo4T<RT> G[kRec];
  const bool has = tid < task.ne;
#pragma unroll
  for (int q = 0; q < kRec; q++)
    if (tid + q * 256 < task.ne) G[q] = push_geo(a, (long long)task.e0 + tid + q * 256);
  // cone inputs: column r of node j of cone cl at buf[(cl * P + j) * R + r]
  if (act) {
    if (a.leaf) {
      // items outer, columns inner: kRC of an item's values load together
      // (each column still sums its items in item order); FP64 takes the
      // columns four at a time to stay within 3 waves' registers
      constexpr int kRC = sizeof(Z) == 16 && R > 4 ? 4 : R;
      const int c = (int)(g - a.leaf_cone_off);
      const int w0 = a.lio[c], w1 = a.lio[c + 1];
#pragma unroll
      for (int r0 = 0; r0 < R; r0 += kRC) {
        Z v[kRC];
#pragma unroll
        for (int r = 0; r < kRC; r++) v[r] = Z{0, 0};
        for (int w = w0; w < w1; w++) {
          Z t[kRC];
#pragma unroll
          for (int r = 0; r < kRC; r++) t[r] = a.lpart[((size_t)w * P + j) * R + r0 + r];
#pragma unroll
          for (int r = 0; r < kRC; r++) v[r] = zadd(v[r], t[r]);
        }
#pragma unroll
        for (int r = 0; r < kRC; r++) buf[(cl * P + j) * R + r0 + r] = v[r];
      }
    } else {
      const int rs = a.rs;
      const int4 ci = a.cidx[g * P + j];
      const Z* s0 = a.pbuf + (size_t)ci.x * rs;
      const Z* s1 = a.pbuf + (size_t)ci.y * rs;
      const Z* s2 = a.pbuf + (size_t)ci.z * rs;
      const Z* s3 = a.pbuf + (size_t)ci.w * rs;
      const Z f = a.ifac_tab[a.ifac_off[g] + j % Ps];
#pragma unroll
      for (int r = 0; r < R; r++)
        buf[(cl * P + j) * R + r] = zmul(zadd(zadd(zadd(s0[r], s1[r]), s2[r]), s3[r]), f);
    }
  }
  __syncthreads();
  if (!a.leaf) {
    // separable transform per column (cone_transform's operations)
    const int n = j / Ps, m = j % Ps;
    if (act) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        Z q{0, 0};
#pragma unroll
        for (int l = 0; l < Pang; l++) {
          const Z f = buf[(cl * P + l * Ps + m) * R + r];
          q.x = fma(f.x, T.Ta[n * 16 + l], q.x);
          q.y = fma(f.y, T.Ta[n * 16 + l], q.y);
        }
        tmp[(cl * P + j) * R + r] = q;
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        Z c{0, 0};
#pragma unroll
        for (int k = 0; k < Ps; k++) {
          const Z v = tmp[(cl * P + n * Ps + k) * R + r];
          c.x = fma(v.x, T.Ts[m * 16 + k], c.x);
          c.y = fma(v.y, T.Ts[m * 16 + k], c.y);
        }
        buf[(cl * P + j) * R + r] = c;
      }
    }
    __syncthreads();
  }
  if constexpr (!kUnitEval) {
    // larger cones: one entry per thread, the coefficients read from LDS
    if (has) {
      RT ts[Ps], ta[Pang];
      ts[0] = 1;
      if (Ps > 1) ts[1] = G[0].s;
#pragma unroll
      for (int i = 2; i < Ps; i++) ts[i] = 2 * G[0].s * ts[i - 1] - ts[i - 2];
      ta[0] = 1;
      if (Pang > 1) ta[1] = G[0].t;
#pragma unroll
      for (int i = 2; i < Pang; i++) ta[i] = 2 * G[0].t * ta[i - 1] - ta[i - 2];
      const Z* c = buf + (size_t)push_entry_cone(a, task, tid) * P * R;
      const Z f{G[0].fre, G[0].fim};
      Z* out = a.pbuf + (size_t)(1 + task.e0 + tid) * a.rs;
#pragma unroll
      for (int r = 0; r < R; r++) {
        Z acc{0, 0};
#pragma unroll
        for (int aa = 0; aa < Pang; aa++) {
          Z row{0, 0};
#pragma unroll
          for (int m = 0; m < Ps; m++) {
            const Z v = c[(aa * Ps + m) * R + r];
            row.x = fma(ts[m], v.x, row.x);
            row.y = fma(ts[m], v.y, row.y);
          }
          acc.x = fma(ta[aa], row.x, acc.x);
          acc.y = fma(ta[aa], row.y, acc.y);
        }
        out[r] = zmul(acc, f);
      }
    }
    return;
  }
  // the records into LDS (tmp's storage; every thread is past the transform)
  Geo4T<RT>* Gs = reinterpret_cast<Geo4T<RT>*>(tmp);
#pragma unroll
  for (int q = 0; q < kRec; q++)
    if (tid + q * 256 < task.ne) Gs[tid + q * 256] = G[q];
  __syncthreads();
  // evaluation units: thread tid = column r of unit u
  const int u = tid / R, r = tid % R;
  const int K = (task.ne + (kUnits - task.nc) - 1) / (kUnits - task.nc);
  int uc = -1, lo = 0, hi = 0, ub = 0;
  for (int q = 0; q < task.nc; q++) {
    const int b0 = max(a.coff[task.c0 + q] - task.e0, 0);
    const int b1 = min(a.coff[task.c0 + q + 1] - task.e0, task.ne);
    const int nu = b1 > b0 ? (b1 - b0 + K - 1) / K : 0;
    if (u >= ub && u < ub + nu) {
      uc = q;
      lo = b0 + (u - ub) * K;
      hi = min(lo + K, b1);
    }
    ub += nu;
  }
  if (uc >= 0) {
    Z cr[P];
#pragma unroll
    for (int q = 0; q < P; q++) cr[q] = buf[((size_t)uc * P + q) * R + r];
    Z* out = a.pbuf + (size_t)(1 + task.e0) * a.rs + r;
    for (int e = lo; e < hi; e++) {
      const Geo4T<RT> g4 = Gs[e];
      // cone_eval's operations
      RT ts[Ps], ta[Pang];
      ts[0] = 1;
      if (Ps > 1) ts[1] = g4.s;
#pragma unroll
      for (int i = 2; i < Ps; i++) ts[i] = 2 * g4.s * ts[i - 1] - ts[i - 2];
      ta[0] = 1;
      if (Pang > 1) ta[1] = g4.t;
#pragma unroll
      for (int i = 2; i < Pang; i++) ta[i] = 2 * g4.t * ta[i - 1] - ta[i - 2];
      Z acc{0, 0};
#pragma unroll
      for (int aa = 0; aa < Pang; aa++) {
        Z row{0, 0};
#pragma unroll
        for (int m = 0; m < Ps; m++) {
          const Z v = cr[aa * Ps + m];
          row.x = fma(ts[m], v.x, row.x);
          row.y = fma(ts[m], v.y, row.y);
        }
        acc.x = fma(ta[aa], row.x, acc.x);
        acc.y = fma(ta[aa], row.y, acc.y);
      }
      out[(size_t)e * a.rs] = zmul(acc, Z{g4.fre, g4.fim});
    }
  }
}

// ---------------------------------------------------------------------------
// Blocks of 8 columns on the matrix cores. The near field of the targets of
// one leaf box (a group) is a dense block V (targets x the box's
// neighbourhood sources, zeros at excluded pairs), and the leaf operators of
// a leaf box's cones form a dense block L ((cone, node) rows x the box's
// sources); both multiply the 8-column block X of the sources' densities. A
// complex product is one real MFMA product:
//   [C_re | C_im] (16 x 16) = [V_re, -V_im] (16 x 2K) [[X_re, X_im], [X_im, -X_re]] (2K x 16),
// source j giving the K rows 2j (V_re / X_re, X_im) and 2j + 1 (-V_im / X_im, -X_re).
// The A operand is stored in MFMA order: tile (row tile rt, K chunk q) holds
// the 64 values lane l reads (row l % 16, K index 4q + l / 16), so a wave
// loads 64 consecutive values per instruction. V_MFMA_F64_16X16X4F64 (FP64
// blocks) and V_MFMA_F32_16X16X4F32 (the mixed-precision blocks of mixed
// plans): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives C[row][l % 16], v < 4, with row = 4 (l / 16)
// + v for the FP32 instruction and row = l / 16 + 4 v for the FP64 one
// (checked on the device with unit-vector operands).
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ inline dbl4 mfma_16x16x4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline flt4 mfma_16x16x4(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// output row of accumulator element v in lane group kk = lane / 16
__device__ inline int mfma_row(double, int kk, int v) { return kk + 4 * v; }
__device__ inline int mfma_row(float, int kk, int v) { return 4 * kk + v; }

// Source index of flat source offset o in a segment table (kMaxNb (end, shift) pairs)
__device__ inline int seg_source(const int* __restrict__ sg, int o) {
  int src = o + sg[1];
#pragma unroll
  for (int q = 1; q < kMaxNb; q++) src = (o >= sg[2 * q - 2]) ? o + sg[2 * q + 1] : src;
  return src;
}

// Build: the A tiles of the near blocks from the stored near-field values.
// One thread per tile value; units give each group's tiles in order.
__global__ void __launch_bounds__(256)
    k_near_mfma_build(long long n, int ngroups, const MfmaGroup* __restrict__ grp, const long long* __restrict__ nf_goff,
                      const zd* __restrict__ nf_val, double* __restrict__ A) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int lo = 0, hi = ngroups - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (grp[mid].aoff <= v) lo = mid;
    else hi = mid - 1;
  }
  const MfmaGroup G = grp[lo];
  const long long local = v - G.aoff;
  const int Q = (G.nsrc + 1) / 2;
  const long long tile = local / 64;
  const int lane = (int)(local % 64);
  const int rt = (int)(tile / Q), q = (int)(tile % Q);
  const int m = lane % 16, kk = lane / 16;
  const int row = rt * 16 + m, o = 2 * q + kk / 2;
  double val = 0.0;
  if (row < G.rows && o < G.nsrc) {
    const int i = G.row0 + row;
    const zd z = nf_val[near_index(nf_goff[i / kNearGroup], i, o)];
    val = (kk % 2 == 0) ? z.x : -z.y;
  }
  A[v] = val;
}

// Build: the A tiles of the leaf blocks from the stored leaf operator (rows
// (cone, node) of the group's cones in order, K over the box's sources).
template <int P, typename Z>
__global__ void __launch_bounds__(256)
    k_leaf_mfma_build(long long n, int ngroups, const MfmaGroup* __restrict__ grp, const int* __restrict__ gcone0,
                      const long long* __restrict__ lf_off, const Z* __restrict__ lf,
                      decltype(Z::x)* __restrict__ A) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int lo = 0, hi = ngroups - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (grp[mid].aoff <= v) lo = mid;
    else hi = mid - 1;
  }
  const MfmaGroup G = grp[lo];
  const long long local = v - G.aoff;
  const int Q = (G.nsrc + 1) / 2;
  const long long tile = local / 64;
  const int lane = (int)(local % 64);
  const int rt = (int)(tile / Q), q = (int)(tile % Q);
  const int m = lane % 16, kk = lane / 16;
  const int row = rt * 16 + m, o = 2 * q + kk / 2;
  decltype(Z::x) val = 0;
  if (row < G.rows && o < G.nsrc) {
    const int c = gcone0[lo] + row / P, node = row % P;
    const Z z = lf[leaf_index<P>(lf_off[c] / P + o, node)];
    val = (kk % 2 == 0) ? z.x : -z.y;
  }
  A[v] = val;
}

// Near field of 8-column blocks: one block of 4 waves per unit (group, row
// tile); wave w sums the K chunks [w Q / 4, (w + 1) Q / 4), the waves' tiles
// are added in wave order; out[(row0 + row) * 8 + r] = the target's near sum
// (RT = double: FP64 tiles and products; float: the tiles rounded to FP32,
// FP32 products and sums, widened into out).
template <typename RT>
__global__ void __launch_bounds__(256)
    k_near_mfma(int nunits, const MfmaUnit* __restrict__ units, const MfmaGroup* __restrict__ grp,
                const int* __restrict__ segs, const RT* __restrict__ A, const zd* __restrict__ rho,
                zd* __restrict__ out) {
  using V4 = typename std::conditional<std::is_same<RT, double>::value, dbl4, flt4>::type;
  __shared__ RT red[3][64][4];
  const int u = xcd_block();
  if (u >= nunits) return;
  const MfmaUnit U = units[u];
  const MfmaGroup G = grp[U.g];
  const int* sg = segs + (size_t)G.seg * kNbSegInts;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int Q = (G.nsrc + 1) / 2;
  const int q0 = w * Q / 4, q1 = (w + 1) * Q / 4;
  const int kk = lane / 16, n = lane % 16;
  const int c = kk % 2, r = n % 8;
  const bool im = n >= 8;
  // the one density component this lane feeds: rows 2j (X_re | X_im), rows
  // 2j + 1 (X_im | -X_re)
  const int part = (c == 0) == im ? 1 : 0;
  const RT sgn = (c == 1 && im) ? RT(-1) : RT(1);
  const double* __restrict__ x = reinterpret_cast<const double*>(rho) + r * 2 + part;
  const RT* __restrict__ a = A + G.aoff + (long long)U.rt * Q * 64 + lane;
  // the group's segment table in registers (uniform)
  int se[kMaxNb], sh[kMaxNb];
#pragma unroll
  for (int q = 0; q < kMaxNb; q++) {
    se[q] = sg[2 * q];
    sh[q] = sg[2 * q + 1];
  }
  // 8 K chunks per round: their loads first (unconditional, at indices
  // clamped into range, so that no load waits on another), then the
  // products on two accumulator chains (added at the end); chunks past the
  // wave's range and sources past the group's take 0
  constexpr int kU = 8;
  V4 acc = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
  for (int qb = q0; qb < q1; qb += kU) {
    RT av[kU], bv[kU];
    bool ok[kU];
#pragma unroll
    for (int t = 0; t < kU; t++) {
      const int q = qb + t < q1 ? qb + t : q1 - 1;
      const int o = 2 * q + kk / 2;
      ok[t] = qb + t < q1 && o < G.nsrc;
      const int oc = o < G.nsrc ? o : G.nsrc - 1;
      int src = oc + sh[0];
#pragma unroll
      for (int k = 1; k < kMaxNb; k++) src = (oc >= se[k - 1]) ? oc + sh[k] : src;
      av[t] = a[(long long)q * 64];
      bv[t] = (RT)x[(size_t)src * 16];
    }
#pragma unroll
    for (int t = 0; t < kU; t++) {
      av[t] = qb + t < q1 ? av[t] : RT(0);
      bv[t] = ok[t] ? sgn * bv[t] : RT(0);
    }
#pragma unroll
    for (int t = 0; t < kU; t += 2) {
      acc = mfma_16x16x4(av[t], bv[t], acc);
      acc2 = mfma_16x16x4(av[t + 1], bv[t + 1], acc2);
    }
  }
  acc += acc2;
  if (w > 0)
    for (int v = 0; v < 4; v++) red[w - 1][lane][v] = acc[v];
  __syncthreads();
  if (w == 0) {
    for (int s = 0; s < 3; s++)
      for (int v = 0; v < 4; v++) acc[v] += red[s][lane][v];
    double* o = reinterpret_cast<double*>(out);
    for (int v = 0; v < 4; v++) {
      const int row = U.rt * 16 + mfma_row(RT(0), kk, v);
      if (row < G.rows) o[((size_t)(G.row0 + row) * 8 + r) * 2 + (im ? 1 : 0)] = (double)acc[v];
    }
  }
}

// Leaf operator of 8-column blocks: one wave per unit (group, row tile), 4
// units per block; out[((cone) * P + node) * 8 + r] = the leaf cone's
// transformed coefficients for column r (row0 = the group's first cone x P).
template <typename RT>
__global__ void __launch_bounds__(256)
    k_leaf_mfma(int nunits, const MfmaUnit* __restrict__ units, const MfmaGroup* __restrict__ grp,
                const RT* __restrict__ A, const zd* __restrict__ rho, typename complex_of<RT>::type* __restrict__ out) {
  using V4 = typename std::conditional<std::is_same<RT, double>::value, dbl4, flt4>::type;
  const int u = xcd_block() * 4 + threadIdx.x / 64;
  if (u >= nunits) return;
  const MfmaUnit U = units[u];
  const MfmaGroup G = grp[U.g];
  const int lane = threadIdx.x % 64;
  const int Q = (G.nsrc + 1) / 2;
  const int kk = lane / 16, n = lane % 16;
  const int c = kk % 2, r = n % 8;
  const bool im = n >= 8;
  const int part = (c == 0) == im ? 1 : 0;
  const RT sgn = (c == 1 && im) ? RT(-1) : RT(1);
  const double* __restrict__ x = reinterpret_cast<const double*>(rho) + r * 2 + part;
  const RT* __restrict__ a = A + G.aoff + (long long)U.rt * Q * 64 + lane;
  constexpr int kU = 8;
  V4 acc = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
  for (int qb = 0; qb < Q; qb += kU) {
    RT av[kU], bv[kU];
    bool ok[kU];
    double xv[kU];
    // unconditional loads at clamped indices, then the zeros (as k_near_mfma)
#pragma unroll
    for (int t = 0; t < kU; t++) {
      const int q = qb + t < Q ? qb + t : Q - 1;
      const int o = 2 * q + kk / 2;
      ok[t] = qb + t < Q && o < G.nsrc;
      av[t] = a[(long long)q * 64];
      xv[t] = x[(size_t)(G.seg + (o < G.nsrc ? o : G.nsrc - 1)) * 16];
    }
#pragma unroll
    for (int t = 0; t < kU; t++) {
      av[t] = qb + t < Q ? av[t] : RT(0);
      bv[t] = ok[t] ? sgn * (RT)xv[t] : RT(0);
    }
#pragma unroll
    for (int t = 0; t < kU; t += 2) {
      acc = mfma_16x16x4(av[t], bv[t], acc);
      acc2 = mfma_16x16x4(av[t + 1], bv[t + 1], acc2);
    }
  }
  acc += acc2;
  RT* o = reinterpret_cast<RT*>(out);
  for (int v = 0; v < 4; v++) {
    const int row = U.rt * 16 + mfma_row(RT(0), kk, v);
    if (row < G.rows) o[((size_t)(G.row0 + row) * 8 + r) * 2 + (im ? 1 : 0)] = acc[v];
  }
}

// k_finalize with R columns: lanes (target, column chunk, slice), S slices
// per (target, chunk), each lane summing its entries for the kFinCols (or R,
// if fewer) columns of its chunk, so a correction weight, an interpolation
// slot index and a near-field partial are loaded once per chunk instead of
// once per column; per column the sums run in the single-column kernel's
// order. xt: the columns transposed in patch order (xt[n * R + r] = x[r * ld
// + n], written by k_gather_b) for the correction's reads; y is column-major
// with stride ld; interpolation entry f holds rs columns at slots[it_slot[f] *
// rs]. stage 0: all of it; stage 1 (beside the level pushes): pre[2 (i R +
// r)] = ident x + correction, pre[2 (i R + r) + 1] = near; stage 2: y = pre[..]
// + (pre[.. + 1] + far), the same bits as stage 0.
constexpr int kFinCols = 4;

template <int S, int R, typename ZI>
__global__ void __launch_bounds__(256)
    k_finalize_b(int N, int Np, double ident, const zd* __restrict__ x, long long ld, const zd* __restrict__ xt,
                 const long long* __restrict__ corr_ptr, const int* __restrict__ corr_patch,
                 const zd* __restrict__ corr_w, const int* __restrict__ inv, const int* __restrict__ nio,
                 const zd* __restrict__ npart, const long long* __restrict__ it_ptr, const int* __restrict__ it_slot,
                 const ZI* __restrict__ slots, int rs, zd* __restrict__ y, int stage = 0,
                 zd* __restrict__ pre = nullptr, const zd* __restrict__ nearb = nullptr) {
  constexpr int CH = R < kFinCols ? R : kFinCols;
  constexpr int NCH = R / CH;
  const long long gt = (long long)xcd_block() * blockDim.x + threadIdx.x;
  const long long ic = gt / S;
  const int sl = (int)(gt % S);
  const int i = (int)(ic / NCH), r0 = (int)(ic % NCH) * CH;
  zd acc[CH], near[CH], far[CH];
#pragma unroll
  for (int c = 0; c < CH; c++) acc[c] = near[c] = far[c] = zd{0.0, 0.0};
  // Each lane sums its entries in the order it, it + S, ...; the loops load
  // kFinBU entries' indices, then their operands, before adding them in that
  // order (one dependent round trip per kFinBU entries, not per entry). The
  // loads are unconditional, at indices clamped into the lane's range, and
  // the entries past it are skipped in the sums: a load inside a conditional
  // would have to complete where the branches meet.
  constexpr int kFinBU = 4;
  if (i < N) {
    const int si = inv[i];
    if (stage != 2) {
      const long long e0 = corr_ptr[i], e1 = corr_ptr[i + 1];
      // (a target's items are few: 32-bit indices within its rows)
      const int ni = (int)((e1 - e0) * Np);
      for (int ib = sl; ib < ni; ib += S * kFinBU) {
        zd wv[kFinBU];
        size_t src[kFinBU];
#pragma unroll
        for (int u = 0; u < kFinBU; u++) {
          const int it0 = ib + u * S;
          const int it = it0 < ni ? it0 : ni - 1;
          const long long e = e0 + it / Np;
          const int q = it % Np;
          wv[u] = corr_w[e * Np + q];
          src[u] = ((size_t)corr_patch[e] * Np + q) * R + r0;
        }
        zd xv[kFinBU][CH];
#pragma unroll
        for (int u = 0; u < kFinBU; u++)
#pragma unroll
          for (int c = 0; c < CH; c++) xv[u][c] = xt[src[u] + c];
#pragma unroll
        for (int u = 0; u < kFinBU; u++)
          if (ib + u * S < ni)
#pragma unroll
            for (int c = 0; c < CH; c++) cfma(acc[c], wv[u], xv[u][c]);
      }
      if (nearb) {
        // the near sum of the 8-column MFMA blocks, one value per (target, column)
        if (sl == 0)
#pragma unroll
          for (int c = 0; c < CH; c++) near[c] = nearb[(size_t)si * R + r0 + c];
      } else {
        const int g = si / kNearGroup, t = si % kNearGroup;
        const int w1 = nio[g + 1];
        for (int wb = nio[g] + sl; wb < w1; wb += S * kFinBU) {
          zd pv[kFinBU][CH];
#pragma unroll
          for (int u = 0; u < kFinBU; u++) {
            const int w = wb + u * S < w1 ? wb + u * S : w1 - 1;
            const zd* np_ = npart + ((size_t)w * kNearGroup + t) * R + r0;
#pragma unroll
            for (int c = 0; c < CH; c++) pv[u][c] = np_[c];
          }
#pragma unroll
          for (int u = 0; u < kFinBU; u++)
            if (wb + u * S < w1)
#pragma unroll
              for (int c = 0; c < CH; c++) near[c] = cadd(near[c], pv[u][c]);
        }
      }
    }
    if (stage != 1) {
      const long long f1 = it_ptr[si + 1];
      for (long long fb = it_ptr[si] + sl; fb < f1; fb += (long long)S * kFinBU) {
        int idx[kFinBU];
#pragma unroll
        for (int u = 0; u < kFinBU; u++) {
          const long long f = fb + (long long)u * S;
          idx[u] = it_slot[f < f1 ? f : f1 - 1];
        }
        ZI sv[kFinBU][CH];
#pragma unroll
        for (int u = 0; u < kFinBU; u++) {
          const ZI* sp = slots + (size_t)idx[u] * rs + r0;
#pragma unroll
          for (int c = 0; c < CH; c++) sv[u][c] = sp[c];
        }
#pragma unroll
        for (int u = 0; u < kFinBU; u++)
          if (fb + (long long)u * S < f1)
#pragma unroll
            for (int c = 0; c < CH; c++) far[c] = cadd(far[c], zwiden(sv[u][c]));
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CH; c++) {
    acc[c] = lane_group_sum<S>(acc[c]);
    near[c] = lane_group_sum<S>(near[c]);
    far[c] = lane_group_sum<S>(far[c]);
  }
  if (i < N && sl == 0) {
#pragma unroll
    for (int c = 0; c < CH; c++) {
      const int r = r0 + c;
      const size_t pi = 2 * ((size_t)i * R + r);
      if (stage == 0) {
        y[(size_t)r * ld + i] = cadd(cadd(cscale(x[(size_t)r * ld + i], ident), acc[c]), cadd(near[c], far[c]));
      } else if (stage == 1) {
        pre[pi] = cadd(cscale(x[(size_t)r * ld + i], ident), acc[c]);
        pre[pi + 1] = near[c];
      } else {
        y[(size_t)r * ld + i] = cadd(pre[pi], cadd(pre[pi + 1], far[c]));
      }
    }
  }
}

}  // namespace ifgf
