// gfx950 kernels of the IFGF forward map (one apply = the six launches below).
//
//   k_gather       rho_s[i] = x[perm[i]] * w[perm[i]]          (SortDensity + compute_intensities)
//   k_near<F>      near field over leaf neighbours             (BoxTree.h:614-683)
//   k_leaf<F,P>    leaf cone values + 2D Chebyshev transform   (BoxTree.h:685-762, Interpolator.h:81-120)
//   k_prop<P>      child -> parent cone propagation            (BoxTree.h:767-868)
//   k_interp<P>    cone interpolation at every target          (BoxTree.h:876-937)
//   k_finalize     y = ident x + correction(x) + unsort(sol_s) (ForwardMap.tpp:481-515,609-655)
//
// Data layout in HBM: points in Morton order as separate x / y / nx / ny FP64
// arrays (coalesced per target thread); cone coefficients of all levels in one
// array [global cone][P] complex, P = Ps*Pang with index itheta*Ps + is
// (Interpolator.h:46-49); per-cone geometry in a 64-byte ConeGeo record.
// Every discrete index (which box, which cone) comes from host tables.
#pragma once

#include <hip/hip_runtime.h>

#include "cone_grid.h"
#include "green.h"
#include "tables.h"

namespace ifgf {

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each with its own L2. Remapping b to a logical block id so that every
// XCD works on one contiguous range of the Morton-ordered work keeps the cone
// coefficients and densities a range touches in that XCD's L2. Bijective on
// [0, gridDim.x).
__device__ inline int xcd_block() {
  constexpr int X = 8;
  const int nb = gridDim.x, b = blockIdx.x;
  const int x = b % X, i = b / X;
  const int base = nb / X, rem = nb % X;
  return x * base + (x < rem ? x : rem) + i;
}

// Two doubles as one 16-byte vector load (a complex value); two floats as one
// 8-byte load (an FP32 complex value).
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt2 __attribute__((ext_vector_type(2)));
template <typename Z>
struct vec2_of;
template <>
struct vec2_of<zd> {
  using type = dbl2;
};
template <>
struct vec2_of<zf> {
  using type = flt2;
};

// Level::Cheb2Cart + Cheb2Radius (Level.h:347-375, 406-411).
__device__ inline void cone_node(const ConeGeo& g, double s, double t, double& x, double& y,
                                 double& r) {
  r = g.h / ((s + 1.0) / 2.0 * g.dr + g.rmin);
  const double th = (t + 1.0) / 2.0 * g.da + g.th0;
  double sn, cs;
  sincos(th, &sn, &cs);
  x = r * cs + g.cx;
  y = r * sn + g.cy;
}

// Level::Cart2Cheb with a given cone + Cheb2Radius (Level.h:424-437, 406-411).
__device__ inline void cone_local(const ConeGeo& g, double x, double y, double& s, double& t,
                                  double& radius) {
  const double dx = x - g.cx, dy = y - g.cy;
  const double r = sqrt(dx * dx + dy * dy);
  const double th = atan2_02pi(dy, dx);
  s = (g.h / r - g.rmin) / g.dr * 2.0 - 1.0;
  t = (th - g.th0) / g.da * 2.0 - 1.0;
  radius = g.h / ((s + 1.0) / 2.0 * g.dr + g.rmin);
}

// Interpolator::Interpolate (Interpolator.h:55-77), in the precision of the
// coefficients (FP64 zd or FP32 zf).
template <int Ps, int Pang, typename Z>
__device__ inline Z cone_eval(const Z* __restrict__ c, decltype(Z::x) s, decltype(Z::x) t) {
  using R = decltype(Z::x);
  R ts[Ps], ta[Pang];
  ts[0] = 1;
  if (Ps > 1) ts[1] = s;
#pragma unroll
  for (int i = 2; i < Ps; i++) ts[i] = 2 * s * ts[i - 1] - ts[i - 2];
  ta[0] = 1;
  if (Pang > 1) ta[1] = t;
#pragma unroll
  for (int i = 2; i < Pang; i++) ta[i] = 2 * t * ta[i - 1] - ta[i - 2];
  Z acc{0, 0};
#pragma unroll
  for (int a = 0; a < Pang; a++) {
    Z row{0, 0};
#pragma unroll
    for (int m = 0; m < Ps; m++) {
      const Z v = c[a * Ps + m];
      row.x = fma(ts[m], v.x, row.x);
      row.y = fma(ts[m], v.y, row.y);
    }
    acc.x = fma(ta[a], row.x, acc.x);
    acc.y = fma(ta[a], row.y, acc.y);
  }
  return acc;
}

// cone_eval at two points of the same cone: every coefficient is read once
// for both (half the LDS reads); each value takes cone_eval's operations in
// cone_eval's order.
template <int Ps, int Pang, typename Z>
__device__ inline void cone_eval2(const Z* __restrict__ c, decltype(Z::x) s0, decltype(Z::x) t0,
                                  decltype(Z::x) s1, decltype(Z::x) t1, Z& v0, Z& v1) {
  using R = decltype(Z::x);
  R ts0[Ps], ta0[Pang], ts1[Ps], ta1[Pang];
  ts0[0] = 1;
  ts1[0] = 1;
  if (Ps > 1) {
    ts0[1] = s0;
    ts1[1] = s1;
  }
#pragma unroll
  for (int i = 2; i < Ps; i++) {
    ts0[i] = 2 * s0 * ts0[i - 1] - ts0[i - 2];
    ts1[i] = 2 * s1 * ts1[i - 1] - ts1[i - 2];
  }
  ta0[0] = 1;
  ta1[0] = 1;
  if (Pang > 1) {
    ta0[1] = t0;
    ta1[1] = t1;
  }
#pragma unroll
  for (int i = 2; i < Pang; i++) {
    ta0[i] = 2 * t0 * ta0[i - 1] - ta0[i - 2];
    ta1[i] = 2 * t1 * ta1[i - 1] - ta1[i - 2];
  }
  Z acc0{0, 0}, acc1{0, 0};
#pragma unroll
  for (int a = 0; a < Pang; a++) {
    Z row0{0, 0}, row1{0, 0};
#pragma unroll
    for (int m = 0; m < Ps; m++) {
      const Z v = c[a * Ps + m];
      row0.x = fma(ts0[m], v.x, row0.x);
      row0.y = fma(ts0[m], v.y, row0.y);
      row1.x = fma(ts1[m], v.x, row1.x);
      row1.y = fma(ts1[m], v.y, row1.y);
    }
    acc0.x = fma(ta0[a], row0.x, acc0.x);
    acc0.y = fma(ta0[a], row0.y, acc0.y);
    acc1.x = fma(ta1[a], row1.x, acc1.x);
    acc1.y = fma(ta1[a], row1.y, acc1.y);
  }
  v0 = acc0;
  v1 = acc1;
}

// Separable 2D Chebyshev transform of the P values of one cone held in LDS
// (Interpolator::GenerateInterpolant, Interpolator.h:81-120). Thread j of the
// cone computes output j; buf and tmp are this cone's P-slot LDS windows.
template <int Ps, int Pang, typename Z, typename TT>
__device__ inline Z cone_transform(const Z* buf, Z* tmp, const TT& T, int j, bool active) {
  const int n = j / Ps, m = j % Ps;
  // stage 1: tmp[n][k] = sum_l f[l][k] Ta[n][l]
  if (active) {
    Z g{0, 0};
#pragma unroll
    for (int l = 0; l < Pang; l++) {
      const Z f = buf[l * Ps + m];
      g.x = fma(f.x, T.Ta[n * 16 + l], g.x);
      g.y = fma(f.y, T.Ta[n * 16 + l], g.y);
    }
    tmp[j] = g;
  }
  __syncthreads();
  // stage 2: C[n][m] = sum_k tmp[n][k] Ts[m][k]
  Z c{0, 0};
  if (active) {
#pragma unroll
    for (int k = 0; k < Ps; k++) {
      const Z v = tmp[n * Ps + k];
      c.x = fma(v.x, T.Ts[m * 16 + k], c.x);
      c.y = fma(v.y, T.Ts[m * 16 + k], c.y);
    }
  }
  return c;
}

// ---------------------------------------------------------------------------
// Every stored-path kernel below is a __device__ body that takes its logical
// block index lb explicitly, and a __global__ wrapper that passes xcd_block()
// (or blockIdx.x): the wrappers run one plan, the fused launches of an
// obstacle group (plan.hip) run every plan's blocks in one grid, each block
// calling the body with its plan's arguments and its block index within the
// plan, so both do the same operations per block.
__device__ inline void gather_body(int lb, int N, const zd* __restrict__ x, const double* __restrict__ w,
                                   const int* __restrict__ perm, zd* __restrict__ rho) {
  const int i = lb * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int o = perm[i];
  rho[i] = cscale(x[o], w[o]);
}
__global__ void __launch_bounds__(256) k_gather(int N, const zd* __restrict__ x,
                                                const double* __restrict__ w,
                                                const int* __restrict__ perm, zd* __restrict__ rho) {
  gather_body(blockIdx.x, N, x, w, perm, rho);
}

// Sum of S lanes' partial results (lanes i*S .. i*S+S-1 of a wave), fixed order.
template <int S, typename R>
__device__ inline R lane_sum(R v) {
#pragma unroll
  for (int o = S / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, S);
  return v;
}
template <int S>
__device__ inline zd lane_group_sum(zd v) {
#pragma unroll
  for (int o = S / 2; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, S);
    v.y += __shfl_xor(v.y, o, S);
  }
  return v;
}

// S lanes per target: lane `sl` takes every S-th source of the neighbour boxes.
template <int F, int S>
__global__ void __launch_bounds__(256)
    k_near(int N, const double* __restrict__ px, const double* __restrict__ py,
           const double* __restrict__ pnx, const double* __restrict__ pny,
           const int* __restrict__ patch, const int4* __restrict__ excl,
           const int* __restrict__ leaf_of_point, const int* __restrict__ leaf_begin,
           const int* __restrict__ nb_ptr, const int* __restrict__ nb_list,
           const zd* __restrict__ rho, zd* __restrict__ sol, double k, double eta) {
  const int lb = xcd_block();
  const int gt = lb * blockDim.x + threadIdx.x;
  const int i = gt / S, sl = gt % S;
  zd acc{0.0, 0.0};
  if (i < N) {
    const double tx = px[i], ty = py[i];
    const int4 e = excl[i];
    const int b = leaf_of_point[i];
    for (int n = nb_ptr[b]; n < nb_ptr[b + 1]; n++) {
      const int nb = nb_list[n];
      const int s1 = leaf_begin[nb + 1];
      for (int s = leaf_begin[nb] + sl; s < s1; s += S) {
        const int ps = patch[s];
        if (ps == e.x || ps == e.y || ps == e.z || ps == e.w) continue;
        const zd g = green_d<F>(tx - px[s], ty - py[s], pnx[s], pny[s], k, eta);
        cfma(acc, g, rho[s]);
      }
    }
  }
  acc = lane_group_sum<S>(acc);
  if (i < N && sl == 0) sol[i] = acc;
}

// Threads (cone, node, slice): S lanes per cone node share the box's sources;
// CPB cones per 256-thread block.
template <int F, int Ps, int Pang, int S>
__global__ void __launch_bounds__(256)
    k_leaf(int ncone, int cone_off, const ConeGeo* __restrict__ cg, const int* __restrict__ cone_box,
           const int* __restrict__ leaf_begin, const double* __restrict__ px,
           const double* __restrict__ py, const double* __restrict__ pnx,
           const double* __restrict__ pny, const zd* __restrict__ rho, zd* __restrict__ coef,
           const ChebTables* __restrict__ tabs, double k, double eta) {
  const int lb = xcd_block();
  constexpr int P = Ps * Pang;
  constexpr int CPB = 256 / (P * S);
  __shared__ zd buf[CPB * P];
  __shared__ zd tmp[CPB * P];
  __shared__ ChebTables T;
  for (int q = threadIdx.x; q < (int)(sizeof(ChebTables) / sizeof(double)); q += blockDim.x)
    ((double*)&T)[q] = ((const double*)tabs)[q];
  __syncthreads();
  const int cl = threadIdx.x / (P * S);
  const int rem = threadIdx.x % (P * S);
  const int j = rem / S, sl = rem % S;
  const int c = lb * CPB + cl;
  const bool active = cl < CPB && c < ncone;
  const int g = cone_off + c;
  zd acc{0.0, 0.0};
  double r = 1.0;
  if (active) {
    const ConeGeo G = cg[g];
    double nx, ny;
    cone_node(G, T.xs[j % Ps], T.xa[j / Ps], nx, ny, r);
    const int b = cone_box[g];
    const int s1 = leaf_begin[b + 1];
    for (int s = leaf_begin[b] + sl; s < s1; s += S) {
      const zd kv = green_d<F>(nx - px[s], ny - py[s], pnx[s], pny[s], k, eta);
      cfma(acc, kv, rho[s]);
    }
  }
  acc = lane_group_sum<S>(acc);
  if (active && sl == 0) buf[cl * P + j] = cmul(acc, inv_centre_factor(r, k));
  __syncthreads();
  // transform: the first CPB*P threads take (cone, node) = (t / P, t % P)
  const int tc = threadIdx.x / P, tj = threadIdx.x % P;
  const bool tact = tc < CPB && lb * CPB + tc < ncone;
  const zd out = cone_transform<Ps, Pang>(buf + (tact ? tc * P : 0), tmp + (tact ? tc * P : 0), T,
                                          tact ? tj : 0, tact);
  if (tact) coef[(size_t)(cone_off + lb * CPB + tc) * P + tj] = out;
}

// ---------------------------------------------------------------------------
// Stored-operator path. The near-field kernel values and the leaf operator
// (kernel -> division by the centre factor -> 2D Chebyshev transform) depend
// only on the geometry; they are evaluated once per plan and each application
// then streams them from HBM.
//
// Near field: target i holds one value per (neighbour box in order, source
// ascending) pair at flat offset o (0 for excluded pairs, same loop order as
// k_near). Storage is grouped for coalescing: the kNearGroup consecutive
// targets of group g = i / kNearGroup are interleaved in chunks of kNearChunk
// offsets, value (i, o) at
//   nf_goff[g] + (o / kNearChunk) * kNearGroup * kNearChunk
//              + (i % kNearGroup) * kNearChunk + o % kNearChunk,
// each target padded with zeros to the group's longest segment (rounded up to
// a chunk). A wave of kNearGroup targets x kNearChunk lanes then reads one
// contiguous kilobyte per load instruction.
constexpr int kNearGroup = 8;
constexpr int kNearChunk = 8;
constexpr int kMaxNb = 9;
constexpr int kNbSegInts = 20;  // 9 (end, shift) pairs padded to 5 int4

__device__ inline long long near_index(long long goff, int i, int o) {
  return goff + (long long)(o / kNearChunk) * (kNearGroup * kNearChunk) + (i % kNearGroup) * kNearChunk +
         o % kNearChunk;
}

// N targets at (qx, qy) (the sorted sources themselves for the forward map);
// leaf_of_point / nb_ptr / nb_list give each target's neighbouring source
// leaf boxes. excl / patch may be null (no excluded pairs: a field at
// separate targets, where coincident points give a zero kernel).
template <int F, int S>
__global__ void __launch_bounds__(256)
    k_near_build(int N, const double* __restrict__ qx, const double* __restrict__ qy,
                 const double* __restrict__ px, const double* __restrict__ py,
                 const double* __restrict__ pnx, const double* __restrict__ pny,
                 const int* __restrict__ patch, const int4* __restrict__ excl,
                 const int* __restrict__ leaf_of_point, const int* __restrict__ leaf_begin,
                 const int* __restrict__ nb_ptr, const int* __restrict__ nb_list,
                 const long long* __restrict__ nf_goff, zd* __restrict__ nf_val, double k, double eta) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt / S, sl = gt % S;
  if (i >= N) return;
  const double tx = qx[i], ty = qy[i];
  const bool ex = excl != nullptr;
  const int4 e = ex ? excl[i] : make_int4(-1, -1, -1, -1);
  const int b = leaf_of_point[i];
  const long long goff = nf_goff[i / kNearGroup];
  int base = 0;
  for (int n = nb_ptr[b]; n < nb_ptr[b + 1]; n++) {
    const int nb = nb_list[n];
    const int s0 = leaf_begin[nb], s1 = leaf_begin[nb + 1];
    for (int s = s0 + sl; s < s1; s += S) {
      const int ps = ex ? patch[s] : 0;
      zd g{0.0, 0.0};
      if (!(ps == e.x || ps == e.y || ps == e.z || ps == e.w))
        g = green_d<F>(tx - px[s], ty - py[s], pnx[s], pny[s], k, eta);
      nf_val[near_index(goff, i, base + (s - s0))] = g;
    }
    base += s1 - s0;
  }
}

// Stored near field in equal work items: one wave per item, an item being
// kNearItemChunks consecutive chunks of one target group (the group's last
// item may hold fewer). Lane (t, sl) of the wave handles target 8g + t and the
// flat offsets o = 8u + sl of the item's chunks u; flat offset o maps to
// source o + shift[q] for the segment q with end[q-1] <= o < end[q], the
// kMaxNb (end, shift) pairs of target i being nf_seg[i] (the last end is the
// row length). Each item writes one partial sum per target of its group to
// npart[item * 8 + t]; k_finalize adds a target's partials in item order.
// Rows vary from a few dozen to several hundred values, and equal items keep
// every wave's dependent chain short (item -> values and segment table ->
// densities), so the stream is not held up by the longest rows.
constexpr int kNearItemChunks = 8;

// kPaced: the form that runs beside the level pushes on the second stream.
// Issuing all of a lane's chunk loads at once made the near stage 41.0 ->
// 38.8 us on the 64k kite but the whole application 1-2% slower there (the
// near field then takes more of the bandwidth the pushes' load phases need);
// on the critical path (one stream, obstacle groups) the unpaced form runs.
template <bool kPaced>
__device__ inline void near_items_body(int lb, int N, int nitems, const NearItem* __restrict__ items,
                                       const int4* __restrict__ nf_seg, const zd* __restrict__ nf_val,
                                       const zd* __restrict__ rho, zd* __restrict__ npart) {
  const int w = lb * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  const int t = lane / kNearChunk, sl = lane % kNearChunk;
  zd acc{0.0, 0.0};
  if (w < nitems) {
    const NearItem it = items[w];
    const int i = it.g * kNearGroup + t;
    if (i < N) {
      const dbl2* __restrict__ v = reinterpret_cast<const dbl2*>(nf_val) + it.off + t * kNearChunk + sl;
      const int4* sg = nf_seg + (size_t)i * (kNbSegInts / 4);
      const int4 q0 = sg[0], q1 = sg[1], q2 = sg[2], q3 = sg[3], q4 = sg[4];
      const int end[kMaxNb - 1] = {q0.x, q0.z, q1.x, q1.z, q2.x, q2.z, q3.x, q3.z};
      const int sh[kMaxNb] = {q0.y, q0.w, q1.y, q1.w, q2.y, q2.w, q3.y, q3.w, q4.y};
      const int len = q4.x;
      dbl2 val[kNearItemChunks];
      if constexpr (!kPaced) {
        // every load unconditional (chunks past the item read its first one,
        // sources past the target's segment read rho[0]) and the entries past
        // the ranges skipped in the sum: all chunks' loads in flight at once
        bool ok[kNearItemChunks];
        zd rv[kNearItemChunks];
#pragma unroll
        for (int u = 0; u < kNearItemChunks; u++) {
          val[u] = v[(u < it.nc ? u : 0) * kNearGroup * kNearChunk];
          const int o = (it.u0 + u) * kNearChunk + sl;
          ok[u] = u < it.nc && o < len;
          int src = o + sh[0];
#pragma unroll
          for (int q = 1; q < kMaxNb; q++) src = (o >= end[q - 1]) ? o + sh[q] : src;
          rv[u] = rho[ok[u] ? src : 0];
        }
#pragma unroll
        for (int u = 0; u < kNearItemChunks; u++)
          if (ok[u]) cfma(acc, zd{val[u].x, val[u].y}, rv[u]);
      } else {
        // loads inside the conditionals: each chunk's density waits for the
        // one before (the same sums in the same order)
#pragma unroll
        for (int u = 0; u < kNearItemChunks; u++)
          if (u < it.nc) val[u] = v[u * kNearGroup * kNearChunk];
#pragma unroll
        for (int u = 0; u < kNearItemChunks; u++) {
          const int o = (it.u0 + u) * kNearChunk + sl;
          if (u < it.nc && o < len) {
            int src = o + sh[0];
#pragma unroll
            for (int q = 1; q < kMaxNb; q++) src = (o >= end[q - 1]) ? o + sh[q] : src;
            cfma(acc, zd{val[u].x, val[u].y}, rho[src]);
          }
        }
      }
    }
  }
  acc = lane_group_sum<kNearChunk>(acc);
  if (w < nitems && sl == 0) npart[(size_t)w * kNearGroup + t] = acc;
}
template <bool kPaced>
__global__ void __launch_bounds__(256)
    k_near_items(int N, int nitems, const NearItem* __restrict__ items, const int4* __restrict__ nf_seg,
                 const zd* __restrict__ nf_val, const zd* __restrict__ rho, zd* __restrict__ npart) {
  near_items_body<kPaced>(xcd_block(), N, nitems, items, nf_seg, nf_val, rho, npart);
}

// Leaf operator L_c (P x n_src per leaf cone c). Storage: the sources of a
// cone are padded to a multiple of kLeafChunk, and the padded columns
// (col = lf_off[c] / P + src, a multiple of kLeafChunk at every cone's start)
// are tiled kLeafChunk at a time: value (col, m) at
//   (col / kLeafChunk) * kLeafChunk * P + m * kLeafChunk + col % kLeafChunk,
// so the lanes (m, src mod kLeafChunk) of the apply read consecutive values.
// Step 1: kernel values divided by the node's centre factor, one thread per
// (cone, source, node) of the padded columns (0 in the padding), written
// column-major to tmp[col * P + j].
constexpr int kLeafChunk = 4;

template <int P>
__device__ inline long long leaf_index(long long col, int m) {
  return (col / kLeafChunk) * (kLeafChunk * P) + m * kLeafChunk + col % kLeafChunk;
}

template <int F, int Ps, int Pang>
__global__ void __launch_bounds__(256)
    k_leaf_build_values(int ncone, int cone_off, const ConeGeo* __restrict__ cg,
                        const int* __restrict__ cone_box, const int* __restrict__ leaf_begin,
                        const long long* __restrict__ lf_off, const long long* __restrict__ item_off,
                        const double* __restrict__ px, const double* __restrict__ py,
                        const double* __restrict__ pnx, const double* __restrict__ pny,
                        const ChebTables* __restrict__ T, zd* __restrict__ tmp, long long nitems,
                        double k, double eta) {
  constexpr int P = Ps * Pang;
  const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems) return;
  // item_off[c] = first item of cone c (items = P * n_src per cone)
  int lo = 0, hi = ncone - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (item_off[mid] <= it) lo = mid;
    else hi = mid - 1;
  }
  const int c = lo;
  const int g = cone_off + c;
  const int b = cone_box[g];
  const int s0 = leaf_begin[b];
  const long long local = it - item_off[c];
  const int j = (int)(local % P);
  const int src = (int)(local / P);
  if (src >= leaf_begin[b + 1] - s0) {
    tmp[lf_off[c] + (long long)src * P + j] = zd{0.0, 0.0};
    return;
  }
  const ConeGeo G = cg[g];
  double nx, ny, r;
  cone_node(G, T->xs[j % Ps], T->xa[j / Ps], nx, ny, r);
  const int s = s0 + src;
  const zd kv = green_d<F>(nx - px[s], ny - py[s], pnx[s], pny[s], k, eta);
  tmp[lf_off[c] + (long long)src * P + j] = cmul(kv, inv_centre_factor(r, k));
}

// Step 2: the 2D Chebyshev transform of every (cone, source) column.
template <int Ps, int Pang>
__global__ void __launch_bounds__(256)
    k_leaf_build_transform(long long ncol, const zd* __restrict__ tmp, const ChebTables* __restrict__ T,
                           zd* __restrict__ lf) {
  constexpr int P = Ps * Pang;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * P) return;
  const long long col = t / P;
  const int m = (int)(t % P);
  const int n = m / Ps, ms = m % Ps;
  const zd* in = tmp + col * P;
  zd acc{0.0, 0.0};
  for (int l = 0; l < Pang; l++) {
    zd row{0.0, 0.0};
    for (int kk = 0; kk < Ps; kk++) {
      const zd v = in[l * Ps + kk];
      row.x = fma(v.x, T->Ts[ms * 16 + kk], row.x);
      row.y = fma(v.y, T->Ts[ms * 16 + kk], row.y);
    }
    acc.x = fma(row.x, T->Ta[n * 16 + l], acc.x);
    acc.y = fma(row.y, T->Ta[n * 16 + l], acc.y);
  }
  lf[leaf_index<P>(col, m)] = acc;
}

// Leaf operator applied in equal work items: an item is up to
// kLeafItemTiles tiles (kLeafChunk sources each) of one leaf cone, threads
// (item, m, slice) with kLeafChunk slices; it writes the partial
// coefficients lpart[item * P + m], which k_level_push sums over the cone's
// items in order. Sources per leaf box range from 1 to over a hundred; equal
// items keep every wave's chain to one item -> values/densities round trip.
constexpr int kLeafItemTiles = 4;

template <int P, typename Z>
__device__ inline void leaf_items_body(int lb, int nitems, const LeafItem* __restrict__ items,
                                       const Z* __restrict__ lf, const zd* __restrict__ rho, Z* __restrict__ lpart) {
  constexpr int S = kLeafChunk;
  using V2 = typename vec2_of<Z>::type;
  const long long gt = (long long)lb * blockDim.x + threadIdx.x;
  const long long item = gt / (P * S);
  const int r = (int)(gt % (P * S));
  const int m = r / S, sl = r % S;
  Z acc{0, 0};
  if (item < nitems) {
    const LeafItem it = items[item];
    const V2* __restrict__ L = reinterpret_cast<const V2*>(lf) + it.off + m * S + sl;
    V2 val[kLeafItemTiles];
    Z rv[kLeafItemTiles];
#pragma unroll
    for (int u = 0; u < kLeafItemTiles; u++) {
      if (u * S + sl < it.n) {
        val[u] = L[u * S * P];
        rv[u] = zcast<Z>(rho[it.s0 + u * S + sl]);
      }
    }
#pragma unroll
    for (int u = 0; u < kLeafItemTiles; u++)
      if (u * S + sl < it.n) zfma(acc, Z{val[u].x, val[u].y}, rv[u]);
  }
  acc.x = lane_sum<S>(acc.x);
  acc.y = lane_sum<S>(acc.y);
  if (item < nitems && sl == 0) lpart[item * P + m] = acc;
}
template <int P, typename Z>
__global__ void __launch_bounds__(256)
    k_leaf_items(int nitems, const LeafItem* __restrict__ items, const Z* __restrict__ lf,
                 const zd* __restrict__ rho, Z* __restrict__ lpart) {
  leaf_items_body<P, Z>(xcd_block(), nitems, items, lf, rho, lpart);
}

// Stored interpolation and propagation geometry: (s, theta, centre factor)
// of every (target, cousin cone) and every (parent node, child cone) pair,
// computed once with the same device functions the on-the-fly kernels use.
template <int S>
__global__ void __launch_bounds__(256)
    k_interp_build(int N, const double* __restrict__ px, const double* __restrict__ py,
                   const long long* __restrict__ it_ptr, const int* __restrict__ it_cone,
                   const ConeGeo* __restrict__ cg, Geo4* __restrict__ it_geo, double k) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt / S, sl = gt % S;
  if (i >= N) return;
  const double x = px[i], y = py[i];
  for (long long e = it_ptr[i] + sl; e < it_ptr[i + 1]; e += S) {
    double s, t, rad;
    cone_local(cg[it_cone[e]], x, y, s, t, rad);
    const zd f = centre_factor(rad, k);
    it_geo[e] = Geo4{s, t, f.x, f.y};
  }
}

// Parent cones of level l-1 (global ids pcone_off..+npc): for node j and the
// k-th relevant child, entry pr_ptr[g] + k*P + j holds the child cone id and
// the node's (s, theta, factor) in it; pn_ifac[g*P + j] = 1/factor of the node.
template <int Ps, int Pang>
__global__ void __launch_bounds__(256)
    k_prop_build(int npc, int pcone_off, int pbox_off, int ccone_off, int cnslot,
                 const ConeGeo* __restrict__ cg, const int* __restrict__ cone_box,
                 const int* __restrict__ cone_q, const int* __restrict__ box_child,
                 const int* __restrict__ proto, const int* __restrict__ dmap,
                 const ChebTables* __restrict__ T, const long long* __restrict__ pr_ptr,
                 int* __restrict__ pr_cone, Geo4* __restrict__ pr_geo, zd* __restrict__ pn_ifac,
                 double k) {
  constexpr int P = Ps * Pang;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npc * P) return;
  const int c = t / P, j = t % P;
  const int g = pcone_off + c;
  double nx, ny, r;
  cone_node(cg[g], T->xs[j % Ps], T->xa[j / Ps], nx, ny, r);
  pn_ifac[(size_t)g * P + j] = inv_centre_factor(r, k);
  const int pb = cone_box[g], q = cone_q[g];
  long long e = pr_ptr[g] + j;
  for (int cp = 0; cp < 4; cp++) {
    const int cb = box_child[(size_t)(pbox_off + pb) * 4 + cp];
    if (cb < 0) continue;
    const int cq = proto[((size_t)q * 4 + cp) * P + j];
    const int cc = ccone_off + dmap[(size_t)cb * cnslot + cq];
    double s, tt, rad;
    cone_local(cg[cc], nx, ny, s, tt, rad);
    const zd f = centre_factor(rad, k);
    pr_cone[e] = cc;
    pr_geo[e] = Geo4{s, tt, f.x, f.y};
    e += P;
  }
}


// ---------------------------------------------------------------------------
// Stored path, one launch per tree level l = D..2 ("push" form of
// BoxTree::Propagation and BoxTree::Interpolation, BoxTree.h:767-937).
//
// Each block owns a task: up to 256 / P cones of level l (or a slice of one
// cone's entries) and a contiguous range of that level's push entries. The
// block first forms the cones' coefficients in LDS:
//   l = D: the sum of the leaf operator items' partial coefficients (already
//          transformed);
//   l < D: node value v_j = (p_0 + p_1 + p_2 + p_3) / factor_j from the
//          values its (up to four) relevant children pushed at the level
//          below, read through cidx (absent children read the zero slot),
//          then the separable 2D transform.
// It then evaluates those interpolants at every push entry (s, theta,
// factor) of the task and writes value * factor to slot 1 + q of pbuf for
// push entry q, so consecutive threads write consecutive slots. A push
// entry is either a parent node at level l-1 falling in the cone
// (propagation) or a (target, cousin cone) pair (interpolation, summed per
// target in fixed order by k_finalize). Every slot has exactly one writer,
// so no atomics and a fixed summation order. The cone of an entry follows
// from the cones' entry offsets coff (entries are grouped by cone).
// The (s, theta, centre factor) of a propagation entry -- parent node j of a
// parent cone with non-relevant id q, evaluated in the cone of its child at
// position cp -- is the same for every parent box (the geometry relative to
// the child's centre depends only on q, cp and j; BoxTree.h:299-428 keys its
// protobox on the same triple), so it is stored once per (level, q, cp, j) in
// ptab; an interpolation entry's is its own, in igeo. einfo tells them apart.
// Likewise the inverse centre factor of parent node j depends only on the
// parent cone's level, radial interval and s node (j % Ps): ifac_tab.
template <typename R>
struct PushArgsT {
  using Z = typename complex_of<R>::type;
  const PushTask* tasks;
  const int* einfo;      // per push entry: >= 0: igeo index; < 0: ptab index -1 - einfo
  const Geo4T<R>* igeo;  // interpolation entries' (s, theta, centre factor), in push order
  const Geo4T<R>* ptab;  // propagation geometry per (level, q, cp) and node
  const int* coff;       // push entries of global cone g: [coff[g], coff[g+1])
  const int4* cidx;      // per (cone, node): slots of its children's pushed values (0: none)
  const int* lio;        // level D: items of leaf cone c: [lio[c], lio[c+1])
  const Z* lpart;        // level D: leaf item partial coefficients [item][P]
  long long leaf_cone_off;
  const int* ifac_off;   // per cone (levels < D): its (level, radial interval) row of ifac_tab
  const Z* ifac_tab;     // 1 / centre factor per (level, radial interval, s node)
  const ChebTablesT<R>* tabs;
  Z* pbuf;               // slot 0 = 0, slot 1 + q = push entry q's value
  int leaf;              // 1 at level D
  int rs;                // batched kernels: columns per slot in pbuf (the widest block)
};

// einfo == nullptr: igeo holds every push entry's record in push order
// (mixed-precision and batched launches, which measured faster without the
// dependent load in front of the geometry).
template <typename R>
__device__ inline Geo4T<R> push_geo(const PushArgsT<R>& a, long long q) {
  if (!a.einfo) return a.igeo[q];
  const int inf = a.einfo[q];
  return inf >= 0 ? a.igeo[inf] : a.ptab[-1 - inf];
}

// Build: the per-entry record of every push entry from the shared tables.
template <typename R>
__global__ void __launch_bounds__(256)
    k_geo_direct(long long n, const int* __restrict__ einfo, const Geo4T<R>* __restrict__ igeo,
                 const Geo4T<R>* __restrict__ ptab, Geo4T<R>* __restrict__ out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int inf = einfo[q];
  out[q] = inf >= 0 ? igeo[inf] : ptab[-1 - inf];
}

// Local cone (0 .. task.nc-1) of entry e of a task: entries are grouped by
// cone in cone order.
template <typename R>
__device__ inline int push_entry_cone(const PushArgsT<R>& a, const PushTask& task, int e) {
  int c = 0;
  for (int q = 1; q < task.nc; q++) c += (e >= a.coff[task.c0 + q] - task.e0) ? 1 : 0;
  return c;
}

// EPT entries per thread: a task holds at most 256 * EPT entries, and every
// thread loads all of its entries together with the cones' inputs before the
// first barrier, so a level costs one memory round trip before the LDS work.
// R = double on a plain plan; float (FP32 coefficients, transforms and
// evaluations, FP64-evaluated centre factors rounded to FP32) on a
// mixed-precision plan.
template <int Ps, int Pang, typename R>
struct PushShared {
  using Z = typename complex_of<R>::type;
  ChebTablesT<R> T;
  Z buf[(256 / (Ps * Pang)) * Ps * Pang];
  Z tmp[(256 / (Ps * Pang)) * Ps * Pang];
};

template <int Ps, int Pang, int EPT, typename R>
__device__ inline void level_push_body(int lb, const PushArgsT<R>& a, PushShared<Ps, Pang, R>& sh) {
  using Z = typename complex_of<R>::type;
  constexpr int P = Ps * Pang;
  ChebTablesT<R>& T = sh.T;
  Z* buf = sh.buf;
  Z* tmp = sh.tmp;
  const PushTask task = a.tasks[lb];
  const int tid = threadIdx.x;
  if (!a.leaf) load_transform_rows<Ps, Pang>(T, a.tabs);
  const int cl = tid / P, j = tid % P;
  const bool act = cl < task.nc;
  const long long g = (long long)task.c0 + cl;
  // entries: EPT >= 2, thread tid takes pairs of consecutive entries
  // 2 (tid + 256 v), + 1 (v < EPT / 2; a task's entries and every cone's run
  // are even in number, so a pair lies in one cone); EPT = 1, entry tid
  constexpr bool kPairs = EPT >= 2;
  auto entry_of = [&](int u) { return kPairs ? 2 * (tid + (u / 2) * 256) + (u & 1) : tid + u * 256; };
  // every thread loads EPT records (indices clamped into the task, so the
  // loads are unconditional and their outstanding count is the same on every
  // path; the extra ones are never used). With entry records (einfo) all EPT
  // record indices are loaded first and each record's address is a select,
  // so the records cost one dependent round trip, not one per entry.
  auto rec_index = [&](int u) -> long long {
    const int e = entry_of(u);
    return (long long)task.e0 + (e < task.ne ? e : task.ne - 1);
  };
  int inf[EPT];
  auto load_info = [&] {
    if (a.einfo) {
#pragma unroll
      for (int u = 0; u < EPT; u++) inf[u] = a.einfo[rec_index(u)];
    }
  };
  Geo4T<R> G[EPT];
  auto load_geo = [&] {
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      const Geo4T<R>* src =
          a.einfo ? (inf[u] >= 0 ? a.igeo + inf[u] : a.ptab + (-1 - inf[u])) : a.igeo + rec_index(u);
      G[u] = *src;
    }
  };
  Z v{0, 0};
  if (a.leaf) {
    load_info();
    load_geo();
    if (act) {
      const int c = (int)(g - a.leaf_cone_off);
      const int w1 = a.lio[c + 1];
      for (int w = a.lio[c]; w < w1; w++) v = zadd(v, a.lpart[(size_t)w * P + j]);
    }
  } else {
    // the children's values first, the entries' records behind them: the
    // records (most of the level's bytes) are still arriving while the
    // block sums the children and runs the transform
    load_info();
    const long long gg = act ? g : (long long)task.c0;
    const int4 ci = a.cidx[gg * P + j];
    const Z fac = a.ifac_tab[a.ifac_off[gg] + j % Ps];
    const Z s0 = a.pbuf[ci.x], s1 = a.pbuf[ci.y], s2 = a.pbuf[ci.z], s3 = a.pbuf[ci.w];
    load_geo();
    v = zmul(zadd(zadd(zadd(s0, s1), s2), s3), fac);
  }
  if (act) buf[tid] = v;
  __syncthreads();
  if (!a.leaf) {
    const Z c = cone_transform<Ps, Pang>(buf + (act ? cl * P : 0), tmp + (act ? cl * P : 0), T, act ? j : 0, act);
    // every thread has passed the transform's barrier, so buf is free again
    if (act) buf[tid] = c;
    __syncthreads();
  }
  if constexpr (kPairs) {
#pragma unroll
    for (int u = 0; u < EPT; u += 2) {
      const int e = entry_of(u);
      if (e < task.ne) {
        const int ce = push_entry_cone(a, task, e);
        Z v0, v1;
        cone_eval2<Ps, Pang>(buf + ce * P, G[u].s, G[u].t, G[u + 1].s, G[u + 1].t, v0, v1);
        a.pbuf[1 + task.e0 + e] = zmul(v0, Z{G[u].fre, G[u].fim});
        a.pbuf[1 + task.e0 + e + 1] = zmul(v1, Z{G[u + 1].fre, G[u + 1].fim});
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      const int e = tid + u * 256;
      if (e < task.ne) {
        const int ce = push_entry_cone(a, task, e);
        const Z val = cone_eval<Ps, Pang>(buf + ce * P, G[u].s, G[u].t);
        a.pbuf[1 + task.e0 + e] = zmul(val, Z{G[u].fre, G[u].fim});
      }
    }
  }
}
template <int Ps, int Pang, int EPT, typename R>
__global__ void __launch_bounds__(256) k_level_push(PushArgsT<R> a) {
  __shared__ PushShared<Ps, Pang, R> sh;
  level_push_body<Ps, Pang, EPT, R>(xcd_block(), a, sh);
}

// FP64 -> FP32 copies of the stored operators of a mixed-precision plan
// (complex values and (s, theta, factor) records alike: n pairs of doubles).
__global__ void __launch_bounds__(256) k_round_f32(long long n, const dbl2* __restrict__ in, flt2* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const dbl2 v = in[i];
  out[i] = flt2{(float)v.x, (float)v.y};
}

// Build: out[i] = in[src[i]] (the push tables gathered from the per-entry
// geometry that k_prop_build and k_interp_build computed).
template <typename T>
__global__ void __launch_bounds__(256)
    k_take(long long n, const long long* __restrict__ src, const T* __restrict__ in, T* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = in[src[i]];
}

// Propagation into the parent cones of level l-1 from the child cones of
// level l. proto: [parent nonrel slot][4][P] child nonrel slot; dmap: level-l
// dense (box, nonrel slot) -> local relevant cone.
template <int Ps, int Pang>
__global__ void __launch_bounds__(256)
    k_prop(int npc, int pcone_off, int pbox_off, int pnrad, int ccone_off, int cnslot,
           const ConeGeo* __restrict__ cg, const int* __restrict__ cone_box,
           const int* __restrict__ cone_q, const int* __restrict__ box_child,
           const int* __restrict__ proto, const int* __restrict__ dmap, zd* __restrict__ coef,
           const ChebTables* __restrict__ tabs, double k) {
  const int lb = xcd_block();
  constexpr int P = Ps * Pang;
  constexpr int CPB = 256 / P;
  __shared__ zd buf[CPB * P];
  __shared__ zd tmp[CPB * P];
  __shared__ ChebTables T;
  for (int q = threadIdx.x; q < (int)(sizeof(ChebTables) / sizeof(double)); q += blockDim.x)
    ((double*)&T)[q] = ((const double*)tabs)[q];
  __syncthreads();
  const int cl = threadIdx.x / P, j = threadIdx.x % P;
  const int c = lb * CPB + cl;
  const bool active = cl < CPB && c < npc;
  const int g = pcone_off + c;
  if (active) {
    const ConeGeo G = cg[g];
    double nx, ny, r;
    cone_node(G, T.xs[j % Ps], T.xa[j / Ps], nx, ny, r);
    const int pb = cone_box[g];
    const int q = cone_q[g];
    zd acc{0.0, 0.0};
#pragma unroll
    for (int cp = 0; cp < 4; cp++) {
      const int cb = box_child[(size_t)(pbox_off + pb) * 4 + cp];
      if (cb < 0) continue;
      const int cq = proto[((size_t)q * 4 + cp) * P + j];
      const int cc = ccone_off + dmap[(size_t)cb * cnslot + cq];
      const ConeGeo C = cg[cc];
      double s, t, rad;
      cone_local(C, nx, ny, s, t, rad);
      const zd v = cone_eval<Ps, Pang>(coef + (size_t)cc * P, s, t);
      cfma(acc, v, centre_factor(rad, k));
    }
    buf[threadIdx.x] = cmul(acc, inv_centre_factor(r, k));
  }
  (void)pnrad;
  __syncthreads();
  const int base = cl * P;
  const zd out = cone_transform<Ps, Pang>(buf + (active ? base : 0), tmp + (active ? base : 0), T,
                                          active ? j : 0, active);
  if (active) coef[(size_t)g * P + j] = out;
}

template <int Ps, int Pang, int S>
__global__ void __launch_bounds__(256)
    k_interp(int N, const double* __restrict__ px, const double* __restrict__ py,
             const long long* __restrict__ it_ptr, const int* __restrict__ it_cone,
             const ConeGeo* __restrict__ cg, const zd* __restrict__ coef, zd* __restrict__ sol,
             double k) {
  const int lb = xcd_block();
  constexpr int P = Ps * Pang;
  const int gt = lb * blockDim.x + threadIdx.x;
  const int i = gt / S, sl = gt % S;
  zd acc{0.0, 0.0};
  if (i < N) {
    const double x = px[i], y = py[i];
    const long long e1 = it_ptr[i + 1];
#pragma unroll 1
    for (long long e = it_ptr[i] + sl; e < e1; e += S) {
      const int g = it_cone[e];
      const ConeGeo C = cg[g];
      double s, t, rad;
      cone_local(C, x, y, s, t, rad);
      const zd v = cone_eval<Ps, Pang>(coef + (size_t)g * P, s, t);
      cfma(acc, v, centre_factor(rad, k));
    }
  }
  acc = lane_group_sum<S>(acc);
  if (i < N && sl == 0) sol[i] = cadd(sol[i], acc);
}

// y[i] = ident x[i] + sum_e W_e x[patch_e] + near[si] + far[si], si = inv[i]
// (a field at separate targets passes x = corr_ptr = null: y[i] = near + far):
// S lanes per target split the correction products, the target's near-field
// item partials (stored path: npart[item * 8 + si % 8] for the items
// [nio[si / 8], nio[si / 8 + 1]) of its group; on the fly: sol[si]) and its
// interpolation entries (stored path: slots[it_slot[f]] for f in
// [it_ptr[si], it_ptr[si+1]), one per (level, cousin cone), in that order);
// lane 0 adds the identity.
// stage 0: all of it. A split in two launches with the same bits: stage 1
// (beside the level pushes, once the near field is done) writes
// pre[2i] = ident x + correction and pre[2i+1] = near; stage 2 adds the
// interpolation entries: y[i] = pre[2i] + (pre[2i+1] + far).
template <int S, typename ZI>
__device__ inline void finalize_body(int lb, int N, int Np, double ident, const zd* __restrict__ x,
                                     const long long* __restrict__ corr_ptr, const int* __restrict__ corr_patch,
                                     const zd* __restrict__ corr_w, const int* __restrict__ inv,
                                     const zd* __restrict__ sol, const int* __restrict__ nio,
                                     const zd* __restrict__ npart, const long long* __restrict__ it_ptr,
                                     const int* __restrict__ it_slot, const ZI* __restrict__ slots,
                                     zd* __restrict__ y, int stage, zd* __restrict__ pre) {
  const int gt = lb * blockDim.x + threadIdx.x;
  const int i = gt / S, sl = gt % S;
  zd acc{0.0, 0.0}, near{0.0, 0.0}, far{0.0, 0.0};
  int si = 0;
  // Each lane walks its entries in the order it, it + S, ... as one sum; the
  // loops below load kFinU entries' indices, then their operands, before
  // adding them in that order, so a lane waits for one dependent round trip
  // per kFinU entries instead of one per entry.
  constexpr int kFinU = 4;
  if (i < N) {
    si = inv[i];
    const bool own = stage != 2;
    const long long e0 = own && corr_ptr ? corr_ptr[i] : 0, e1 = own && corr_ptr ? corr_ptr[i + 1] : 0;
    const long long nitem = (e1 - e0) * Np;
    // (the loads are unconditional, at indices clamped into the lane's
    // range, and the entries past it are skipped in the sums: a load inside a
    // conditional would have to complete where the branches meet)
    // (a target's items are few: 32-bit indices within its rows)
    const int ni = (int)nitem;
    for (int ib = sl; ib < ni; ib += S * kFinU) {
      zd wv[kFinU], xv[kFinU];
      long long src[kFinU];
#pragma unroll
      for (int u = 0; u < kFinU; u++) {
        const int it0 = ib + u * S;
        const int it = it0 < ni ? it0 : ni - 1;
        const long long e = e0 + it / Np;
        const int q = it % Np;
        wv[u] = corr_w[e * Np + q];
        src[u] = (long long)corr_patch[e] * Np + q;
      }
#pragma unroll
      for (int u = 0; u < kFinU; u++) xv[u] = x[src[u]];
#pragma unroll
      for (int u = 0; u < kFinU; u++)
        if (ib + u * S < ni) cfma(acc, wv[u], xv[u]);
    }
    if (!own) {
    } else if (npart) {
      const int g = si / kNearGroup, t = si % kNearGroup;
      const int w1 = nio[g + 1];
      for (int wb = nio[g] + sl; wb < w1; wb += S * kFinU) {
        zd pv[kFinU];
#pragma unroll
        for (int u = 0; u < kFinU; u++)
          pv[u] = npart[(size_t)(wb + u * S < w1 ? wb + u * S : w1 - 1) * kNearGroup + t];
#pragma unroll
        for (int u = 0; u < kFinU; u++)
          if (wb + u * S < w1) near = cadd(near, pv[u]);
      }
    } else if (sl == 0) {
      near = sol[si];
    }
    if (slots && stage != 1) {
      const long long f1 = it_ptr[si + 1];
      for (long long fb = it_ptr[si] + sl; fb < f1; fb += (long long)S * kFinU) {
        int idx[kFinU];
#pragma unroll
        for (int u = 0; u < kFinU; u++) {
          const long long f = fb + (long long)u * S;
          idx[u] = it_slot[f < f1 ? f : f1 - 1];
        }
        ZI sv[kFinU];
#pragma unroll
        for (int u = 0; u < kFinU; u++) sv[u] = slots[idx[u]];
#pragma unroll
        for (int u = 0; u < kFinU; u++)
          if (fb + (long long)u * S < f1) far = cadd(far, zwiden(sv[u]));
      }
    }
  }
  acc = lane_group_sum<S>(acc);
  near = lane_group_sum<S>(near);
  far = lane_group_sum<S>(far);
  if (i < N && sl == 0) {
    if (stage == 0) {
      y[i] = cadd(cadd(x ? cscale(x[i], ident) : zd{0.0, 0.0}, acc), cadd(near, far));
    } else if (stage == 1) {
      pre[2 * (size_t)i] = cadd(x ? cscale(x[i], ident) : zd{0.0, 0.0}, acc);
      pre[2 * (size_t)i + 1] = near;
    } else {
      y[i] = cadd(pre[2 * (size_t)i], cadd(pre[2 * (size_t)i + 1], far));
    }
  }
}
template <int S, typename ZI>
__global__ void __launch_bounds__(256)
    k_finalize(int N, int Np, double ident, const zd* __restrict__ x,
               const long long* __restrict__ corr_ptr, const int* __restrict__ corr_patch,
               const zd* __restrict__ corr_w, const int* __restrict__ inv, const zd* __restrict__ sol,
               const int* __restrict__ nio, const zd* __restrict__ npart,
               const long long* __restrict__ it_ptr, const int* __restrict__ it_slot, const ZI* __restrict__ slots,
               zd* __restrict__ y, int stage = 0, zd* __restrict__ pre = nullptr) {
  finalize_body<S, ZI>(xcd_block(), N, Np, ident, x, corr_ptr, corr_patch, corr_w, inv, sol, nio, npart, it_ptr,
                       it_slot, slots, y, stage, pre);
}

// Direct sum at a target subset, excluding the same pairs the correction
// replaces (compute_non_singular_interactions_unacc, ForwardMap.tpp:438-478);
// excl_t holds kMaxExcl excluded patches per target.
template <int F>
__global__ void __launch_bounds__(256)
    k_direct(int nt, const int* __restrict__ targets, const int* __restrict__ excl_t, int N, int Np,
             const double* __restrict__ ox, const double* __restrict__ oy,
             const double* __restrict__ onx, const double* __restrict__ ony,
             const zd* __restrict__ rho_o, zd* __restrict__ out, double k, double eta) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int i = targets[t];
  const int* e = excl_t + (size_t)t * 4;
  const double tx = ox[i], ty = oy[i];
  zd acc{0.0, 0.0};
  for (int s = 0; s < N; s++) {
    const int ps = s / Np;
    if (ps == e[0] || ps == e[1] || ps == e[2] || ps == e[3]) continue;
    cfma(acc, green<F>(tx, ty, ox[s], oy[s], onx[s], ony[s], k, eta), rho_o[s]);
  }
  out[t] = acc;
}

// Off-surface field (propagate_unacc, ForwardMap.tpp:702-727).
template <int F>
__global__ void __launch_bounds__(256)
    k_propagate(int nt, const double* __restrict__ tx, const double* __restrict__ ty, int N,
                const double* __restrict__ ox, const double* __restrict__ oy,
                const double* __restrict__ onx, const double* __restrict__ ony,
                const zd* __restrict__ rho_o, zd* __restrict__ out, double k, double eta) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  zd acc{0.0, 0.0};
  for (int s = 0; s < N; s++) cfma(acc, green<F>(tx[t], ty[t], ox[s], oy[s], onx[s], ony[s], k, eta), rho_o[s]);
  out[t] = acc;
}

__global__ void __launch_bounds__(256)
    k_hankel(long long n, const double* __restrict__ x, zd* __restrict__ h0, zd* __restrict__ h1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  zd a, b;
  hankel01(x[i], a, b);
  h0[i] = a;
  h1[i] = b;
}

}  // namespace ifgf
