// Tree setup on the GPU: the grouping step (each target's box at every level,
// Level::Point2Morton, IFGF_Source/Level.h:145-161; the cousin lists of
// Level::GetCousins, :261-289, and the neighbourhoods of GetNeighbours,
// :378-404; the interpolation entry counts) and the cone step
// (BoxTree::InitializeRelevantConeSegments, BoxTree.h:553-583, with
// GetRelevantConeSegmentsDueToParents, :299-428, the merge of :430-459 and
// the flattening of :461-514), producing the host loops' tables.
//
//   k_target_boxes    one thread per target: its box code at every level and,
//                     for the forward map, the box's index (binary search in
//                     the level's ascending box codes)
//   k_group_runs      one workgroup per level of a field: the runs of equal
//                     codes among the sorted targets are its groups
//   k_window          per group: the boxes of its cousin window (6 x 6 less
//                     the 3 x 3 neighbourhood) or of its 3 x 3 neighbourhood
//                     that exist, counted, then listed in ascending order
//   k_scan            one workgroup: exclusive prefix sums with a carry
//   k_entry_counts    per target: its interpolation entries over the levels
//
//   k_interp_entries  one thread per target: the non-relevant cone of each of
//                     its (level, cousin box) interpolation entries, marked
//   k_proto_first     per (parent direction q, child position cp): the first
//                     relevant parent cone with that direction and that child
//   k_proto_fill      that cone's interpolation nodes located in the child's
//                     cone grid (the level's protobox)
//   k_proto_mark      the child cones every parent cone's nodes fall in
//   k_number_level    one workgroup per level: relevant cones numbered box by
//                     box, direction ascending (a block scan with a carry)
//   k_it_cone         each interpolation entry's global cone id
//
// The geometry goes through the functions the host loop calls (cone_grid.h)
// and the numbering follows the host's order. A point within a relative 1e-9
// of a cone boundary (parent nodes on the diagonal through a parent's and a
// child's centre land exactly on a child direction edge) is not decided on the
// GPU: its cone is computed on the host with the host's atan, cos and sin,
// whose last bit decides it there. So the tables are the host's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "cone_grid.h"
#include "plan_internal.h"

namespace ifgf {
namespace {

constexpr int kThreads = 256;
constexpr int kNumberThreads = 1024;

// Interpolation entry e (target i, level l, cousin box c) and protobox node
// idx (parent box pb, child box cb) left to the host.
struct NearEntry {
  long long e;
  int i, l, c, pad;
};
struct NearNode {
  long long idx;
  int pb, cb;
};
constexpr int kNearCap = 1 << 20;

struct LevelDev {
  double min_x, min_y, bs, h;
  int nce, nrad, radial_off, nslot;
  int box_off, cone_off, pad0, pad1;
  long long map_off;
};

__device__ inline ConeGrid grid_of(const LevelDev& L, const double* radial) {
  return ConeGrid{L.min_x, L.min_y, L.bs, L.h, L.nce, L.nrad, radial + L.radial_off};
}

inline unsigned grid_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

__global__ void __launch_bounds__(kThreads) k_fill(long long n, int v, int* __restrict__ a) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) a[i] = v;
}

// flag bits: 1 target outside a cousin's cone grid, 2 parent node outside the
// child's cone grid; nnear counts the entries / nodes left to the host (only
// the first kNearCap are recorded)
__global__ void __launch_bounds__(kThreads)
    k_interp_entries(long long nt, int D, const double* __restrict__ tx, const double* __restrict__ ty,
                     const int* __restrict__ pbox, const long long* __restrict__ grp_off,
                     const int* __restrict__ cz_ptr, const int* __restrict__ cz_list,
                     const long long* __restrict__ it_ptr, const LevelDev* __restrict__ lv,
                     const double* __restrict__ radial, const long long* __restrict__ box_morton,
                     int* __restrict__ ent, int* __restrict__ mark, int* __restrict__ flag,
                     NearEntry* __restrict__ near, int* __restrict__ nnear) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nt) return;
  const double x = tx[i], y = ty[i];
  long long e = it_ptr[i];
  for (int l = D; l >= 2; l--) {
    const LevelDev L = lv[l];
    const ConeGrid G = grid_of(L, radial);
    const long long g = grp_off[l] + pbox[i * (D + 1) + l];
    for (int n = cz_ptr[g]; n < cz_ptr[g + 1]; n++) {
      const int c = cz_list[n];
      bool nr;
      const long long q = grid_nonrel_near(G, box_morton[L.box_off + c], x, y, nr);
      if (nr) {
        const int k = atomicAdd(nnear, 1);
        if (k < kNearCap) near[k] = NearEntry{e, (int)i, l, c, 0};
        ent[e++] = -1;
        continue;
      }
      if (q < 0) {
        atomicOr(flag, 1);
        ent[e++] = -1;
        continue;
      }
      const long long a = L.map_off + (long long)c * L.nslot + q;
      ent[e++] = (int)a;
      mark[a] = 1;
    }
  }
}

__global__ void __launch_bounds__(kThreads)
    k_proto_first(int ncone, int cone_off, int nrad, int box_off, const int* __restrict__ cone_box,
                  const int* __restrict__ cone_posr, const int* __restrict__ cone_post,
                  const int* __restrict__ box_child, int* __restrict__ first) {
  const int gl = blockIdx.x * kThreads + threadIdx.x;
  if (gl >= ncone) return;
  const int g = cone_off + gl, pb = cone_box[g];
  const long long q = (long long)cone_post[g] * nrad + cone_posr[g];
  for (int cp = 0; cp < 4; cp++)
    if (box_child[(size_t)(box_off + pb) * 4 + cp] >= 0) atomicMin(&first[q * 4 + cp], g);
}

// proto[(q * 4 + cp) * P + j] for the (q, cp) that have a first parent cone;
// the others keep -1
__global__ void __launch_bounds__(kThreads)
    k_proto_fill(long long n, int P, int Ps, int l, const LevelDev* __restrict__ lv, const double* __restrict__ radial,
                 const long long* __restrict__ box_morton, const int* __restrict__ cone_box,
                 const int* __restrict__ box_child, const int* __restrict__ first, const double* __restrict__ xs,
                 const double* __restrict__ xa, int* __restrict__ proto, int* __restrict__ flag,
                 NearNode* __restrict__ near, int* __restrict__ nnear) {
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= n) return;
  const long long qc = idx / P;
  const int j = (int)(idx % P);
  const int g = first[qc];
  if (g == INT_MAX) return;
  const LevelDev Pl = lv[l - 1], L = lv[l];
  const long long q = qc / 4;
  const int cp = (int)(qc % 4), pb = cone_box[g];
  const int cb = box_child[(size_t)(Pl.box_off + pb) * 4 + cp];
  double px, py;
  grid_cheb2cart(grid_of(Pl, radial), box_morton[Pl.box_off + pb], q, xs[j % Ps], xa[j / Ps], px, py);
  bool nr;
  const long long cq = grid_nonrel_near(grid_of(L, radial), box_morton[L.box_off + cb], px, py, nr);
  if (nr) {
    const int k = atomicAdd(nnear, 1);
    if (k < kNearCap) near[k] = NearNode{idx, pb, cb};
    return;
  }
  if (cq < 0) {
    atomicOr(flag, 2);
    return;
  }
  proto[idx] = (int)cq;
}

__global__ void __launch_bounds__(kThreads)
    k_proto_mark(long long n, int P, int l, const LevelDev* __restrict__ lv, const int* __restrict__ cone_box,
                 const int* __restrict__ cone_posr, const int* __restrict__ cone_post,
                 const int* __restrict__ box_child, const int* __restrict__ proto, int* __restrict__ mark,
                 unsigned long long* __restrict__ nprop) {
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= n) return;
  const LevelDev Pl = lv[l - 1], L = lv[l];
  const int gl = (int)(idx / (4 * P)), r = (int)(idx % (4 * P)), cp = r / P, j = r % P;
  const int g = Pl.cone_off + gl, pb = cone_box[g];
  const int cb = box_child[(size_t)(Pl.box_off + pb) * 4 + cp];
  if (cb < 0) return;
  const long long q = (long long)cone_post[g] * Pl.nrad + cone_posr[g];
  const int cq = proto[(q * 4 + cp) * P + j];
  if (cq < 0) return;  // only after a flagged error
  mark[L.map_off + (long long)cb * L.nslot + cq] = 1;
  if (j == 0) atomicAdd(nprop, (unsigned long long)P);
}

// One workgroup numbers the marked (box, direction) slots of one level in
// ascending order: local id -> dense_map, global id cone_off + local -> box,
// posr, post; unmarked slots get -1. count = the level's relevant cones.
__global__ void __launch_bounds__(kNumberThreads)
    k_number_level(long long map_off, long long n, int nslot, int nrad, int cone_off, const int* __restrict__ mark,
                   int* __restrict__ dense_map, int* __restrict__ cone_box, int* __restrict__ cone_posr,
                   int* __restrict__ cone_post, int* __restrict__ count) {
  constexpr int kW = kNumberThreads / 64;
  __shared__ int wsum[kW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int carry = 0;
  for (long long base = 0; base < n; base += kNumberThreads) {
    const long long a = base + threadIdx.x;
    const int m = a < n ? (mark[map_off + a] != 0) : 0;
    const unsigned long long bal = __ballot(m);
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int k = 0; k < kW; k++) {
      woff += k < w ? wsum[k] : 0;
      tot += wsum[k];
    }
    if (a < n) {
      if (m) {
        const int id = carry + woff + pre, g = cone_off + id;
        const int q = (int)(a % nslot);
        dense_map[map_off + a] = id;
        cone_box[g] = (int)(a / nslot);
        cone_posr[g] = q % nrad;
        cone_post[g] = q / nrad;
      } else {
        dense_map[map_off + a] = -1;
      }
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = carry;
}

// host decisions: dst[pos[k]] = val[k] (and mark[val[k]] = 1 when mark is given)
__global__ void __launch_bounds__(kThreads)
    k_put(int n, const long long* __restrict__ pos, const int* __restrict__ val, int* __restrict__ dst,
          int* __restrict__ mark) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  dst[pos[k]] = val[k];
  if (mark) mark[val[k]] = 1;
}

__global__ void __launch_bounds__(kThreads)
    k_it_cone(long long nent, int D, const LevelDev* __restrict__ lv, const int* __restrict__ ent,
              const int* __restrict__ dense_map, int* __restrict__ it_cone) {
  const long long e = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (e >= nent) return;
  const int a = ent[e];
  int l = 2;
  for (int k = 3; k <= D; k++)
    if (a >= lv[k].map_off) l = k;
  it_cone[e] = lv[l].cone_off + dense_map[a];
}

// ------------------------------------------------------------- grouping ---
struct GLevel {
  double min_x, min_y, bs;
  long long nbps;
  long long grp_off;  // first global group id of the level
  int box_off, nbox;  // the level's boxes in the global box arrays
  int ngrp, pad;
};

// index of code in the ascending codes c[0..n), or -1
__device__ inline int find_code(const long long* __restrict__ c, int n, long long code) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] < code)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && c[lo] == code ? lo : -1;
}

// flag bits: 1 a target outside the level's box grid (field), 2 a point
// outside the relevant boxes of its level (forward map)
__global__ void __launch_bounds__(kThreads)
    k_target_boxes(long long nt, int D, int sep, const double* __restrict__ tx, const double* __restrict__ ty,
                   const GLevel* __restrict__ glv, const long long* __restrict__ box_morton,
                   long long* __restrict__ codes, int* __restrict__ pbox, int* __restrict__ flag) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nt) return;
  const double x = tx[i], y = ty[i];
  pbox[i * (D + 1)] = -1;
  pbox[i * (D + 1) + 1] = -1;
  for (int l = 2; l <= D; l++) {
    const GLevel L = glv[l];
    const long long px = (long long)((x - L.min_x) / L.bs), py = (long long)((y - L.min_y) / L.bs);
    const bool ok = !(px < 0 || py < 0 || px >= L.nbps || py >= L.nbps);
    const long long m = ok ? morton_code(px, py, l) : -1;
    codes[(size_t)l * nt + i] = m;
    if (sep) {
      if (!ok) atomicOr(flag, 1);
      continue;
    }
    const int b = ok ? find_code(box_morton + L.box_off, L.nbox, m) : -1;
    if (b < 0) atomicOr(flag, 2);
    pbox[i * (D + 1) + l] = b;
  }
}

// Field level l: group of sorted target i = number of code changes before
// it; gcode[g] = the group's code; count = groups.
__global__ void __launch_bounds__(kNumberThreads)
    k_group_runs(long long nt, int D, int l, const long long* __restrict__ code, int* __restrict__ pbox,
                 long long* __restrict__ gcode, int* __restrict__ count) {
  constexpr int kW = kNumberThreads / 64;
  __shared__ int wsum[kW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int carry = 0;
  for (long long base = 0; base < nt; base += kNumberThreads) {
    const long long i = base + threadIdx.x;
    const int f = i < nt ? (i == 0 || code[i] != code[i - 1]) : 0;
    const unsigned long long bal = __ballot(f);
    const int pre = __popcll(bal & ((2ull << lane) - 1ull));  // inclusive
    if (lane == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int k = 0; k < kW; k++) {
      woff += k < w ? wsum[k] : 0;
      tot += wsum[k];
    }
    if (i < nt) {
      const int g = carry + woff + pre - 1;
      pbox[i * (D + 1) + l] = g;
      if (f) gcode[g] = code[i];
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = carry;
}

// Groups [g0, g0 + n) of one level, codes gcode: the existing boxes of the
// cousin window (cousins = 1) or of the 3 x 3 neighbourhood (cousins = 0).
// out == nullptr: cnt[g] = their number; else out[ptr[g] ..] = their local
// box indices, ascending (ascending codes).
__global__ void __launch_bounds__(kThreads)
    k_window(int n, int l, int cousins, const long long* __restrict__ gcode, const GLevel* __restrict__ glv,
             const long long* __restrict__ box_morton, int* __restrict__ cnt, const int* __restrict__ ptr,
             int* __restrict__ out) {
  const int g = blockIdx.x * kThreads + threadIdx.x;
  if (g >= n) return;
  const GLevel L = glv[l];
  const long long* codes = box_morton + L.box_off;
  const long long m = gcode[g];
  const long long bx = (long long)morton_merge2((uint64_t)m), by = (long long)morton_merge2((uint64_t)m >> 1);
  long long xlo, xup, ylo, yup;
  if (cousins) {
    xlo = (bx / 2) * 2 - 2 > 0 ? (bx / 2) * 2 - 2 : 0;
    xup = (bx / 2) * 2 + 3 < L.nbps - 1 ? (bx / 2) * 2 + 3 : L.nbps - 1;
    ylo = (by / 2) * 2 - 2 > 0 ? (by / 2) * 2 - 2 : 0;
    yup = (by / 2) * 2 + 3 < L.nbps - 1 ? (by / 2) * 2 + 3 : L.nbps - 1;
  } else {
    xlo = bx - 1 > 0 ? bx - 1 : 0;
    xup = bx + 1 < L.nbps - 1 ? bx + 1 : L.nbps - 1;
    ylo = by - 1 > 0 ? by - 1 : 0;
    yup = by + 1 < L.nbps - 1 ? by + 1 : L.nbps - 1;
  }
  int found[36];
  int k = 0;
  for (long long j = ylo; j <= yup; j++)
    for (long long i = xlo; i <= xup; i++) {
      if (cousins && (i - bx <= 1 && bx - i <= 1) && (j - by <= 1 && by - j <= 1)) continue;
      const int b = find_code(codes, L.nbox, morton_code(i, j, l));
      if (b >= 0 && k < 36) found[k++] = b;
    }
  if (!out) {
    cnt[g] = k;
    return;
  }
  for (int a = 1; a < k; a++) {  // insertion sort, at most 27 entries
    const int v = found[a];
    int b = a - 1;
    while (b >= 0 && found[b] > v) {
      found[b + 1] = found[b];
      b--;
    }
    found[b + 1] = v;
  }
  int* o = out + ptr[g];
  for (int a = 0; a < k; a++) o[a] = found[a];
}

// out[i] = sum of in[0..i) (exclusive), out[n] = the total; one workgroup.
template <typename T>
__global__ void __launch_bounds__(kNumberThreads) k_scan(long long n, const int* __restrict__ in, T* __restrict__ out) {
  constexpr int kW = kNumberThreads / 64;
  __shared__ T wsum[kW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T carry = 0;
  for (long long base = 0; base < n; base += kNumberThreads) {
    const long long i = base + threadIdx.x;
    const T v = i < n ? (T)in[i] : (T)0;
    T x = v;  // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
      const T y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    T woff = 0, tot = 0;
    for (int k = 0; k < kW; k++) {
      woff += k < w ? wsum[k] : (T)0;
      tot += wsum[k];
    }
    if (i < n) out[i] = carry + woff + x - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ void __launch_bounds__(kThreads)
    k_entry_counts(long long nt, int D, const int* __restrict__ pbox, const GLevel* __restrict__ glv,
                   const int* __restrict__ cz_ptr, int* __restrict__ cnt) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nt) return;
  int c = 0;
  for (int l = 2; l <= D; l++) {
    const long long g = glv[l].grp_off + pbox[i * (D + 1) + l];
    c += cz_ptr[g + 1] - cz_ptr[g];
  }
  cnt[i] = c;
}

template <typename T>
void download(std::vector<T>& v, const DevBuf<T>& b, size_t n) {
  v.resize(n);
  if (n) HIP_TRY(hipMemcpy(v.data(), b.p, n * sizeof(T), hipMemcpyDeviceToHost));
}

template <typename T>
T read_one(const T* p) {
  T v;
  HIP_TRY(hipMemcpy(&v, p, sizeof(T), hipMemcpyDeviceToHost));
  return v;
}

}  // namespace

bool relevant_cones_device(const ConeInput& in, Tree& t, std::string& err) {
  const int D = t.D, P = t.Ps * t.Pang;
  const long long nent = t.it_ptr[in.nt];
  long long map_total = 0, proto_total = 0;
  for (int l = 2; l <= D; l++) {
    map_total = t.lv[l].map_off + (long long)t.lv[l].nbox * t.lv[l].nslot;
    if (l > 2) proto_total = t.lv[l].proto_off + (long long)t.lv[l - 1].nslot * 4 * P;
  }
  if (map_total > 2147483647LL) {
    err = "cone map too large";
    return false;
  }
  std::vector<LevelDev> lvh(D + 1);
  for (int l = 2; l <= D; l++) {
    const LevelInfo& L = t.lv[l];
    LevelDev& d = lvh[l];
    d = LevelDev{};
    d.min_x = L.min_x;
    d.min_y = L.min_y;
    d.bs = L.bs;
    d.h = L.h;
    d.nce = L.nce;
    d.nrad = L.nrad;
    d.radial_off = L.radial_off;
    d.nslot = L.nslot;
    d.box_off = L.box_off;
    d.map_off = L.map_off;
  }
  std::vector<long long> nt_ptr(t.it_ptr.begin(), t.it_ptr.begin() + in.nt + 1);
  DevBuf<LevelDev> lv;
  DevBuf<double> radial, tx, ty, xs, xa;
  DevBuf<long long> box_morton, grp_off, it_ptr;
  DevBuf<int> box_child, pbox, cz_ptr, cz_list, ent, mark, dense_map, cone_box, cone_posr, cone_post, proto, first,
      it_cone, scal;
  DevBuf<unsigned long long> nprop;
  DevBuf<NearEntry> near_e;
  DevBuf<NearNode> near_n;
  near_e.alloc(std::min<long long>(nent, kNearCap));
  near_n.alloc(std::min<long long>(proto_total, kNearCap));
  lv.upload(lvh);
  radial.upload(t.radial);
  box_morton.upload(t.box_morton);
  box_child.upload(t.box_child);
  tx.upload(std::vector<double>(in.tx, in.tx + in.nt));
  ty.upload(std::vector<double>(in.ty, in.ty + in.nt));
  pbox.upload(std::vector<int>(in.pbox, in.pbox + (size_t)in.nt * (D + 1)));
  grp_off.upload(in.grp_off);
  cz_ptr.upload(in.cz_ptr);
  cz_list.upload(in.cz_list);
  it_ptr.upload(nt_ptr);
  xs.upload(in.xs);
  xa.upload(in.xa);
  ent.alloc(nent);
  mark.alloc(map_total);
  dense_map.alloc(map_total);
  // a level has at most nbox * nslot relevant cones
  cone_box.alloc(map_total);
  cone_posr.alloc(map_total);
  cone_post.alloc(map_total);
  proto.alloc(proto_total);
  scal.alloc(3 + (size_t)D + 1);  // [0] error flags, [1] / [2] near entries / nodes, [3 + l] cones of level l
  nprop.alloc(1);
  HIP_TRY(hipMemset(mark.p, 0, map_total * sizeof(int)));
  HIP_TRY(hipMemset(scal.p, 0, scal.n * sizeof(int)));
  HIP_TRY(hipMemset(nprop.p, 0, sizeof(unsigned long long)));
  if (proto_total) HIP_TRY(hipMemset(proto.p, 0xff, proto_total * sizeof(int)));  // -1
  if (in.nt)
    hipLaunchKernelGGL(k_interp_entries, dim3(grid_for(in.nt)), dim3(kThreads), 0, 0, in.nt, D, tx.p, ty.p, pbox.p,
                       grp_off.p, cz_ptr.p, cz_list.p, it_ptr.p, lv.p, radial.p, box_morton.p, ent.p, mark.p, scal.p,
                       near_e.p, scal.p + 1);
  HIP_TRY(hipGetLastError());
  if (read_one(scal.p) & 1) {
    err = "target outside the cone grid of a cousin box";
    return false;
  }
  auto host_grid = [&](int l) {
    const LevelInfo& L = t.lv[l];
    return ConeGrid{L.min_x, L.min_y, L.bs, L.h, L.nce, L.nrad, t.radial.data() + L.radial_off};
  };
  // hand the host's decisions back: dst[pos[k]] = val[k]
  DevBuf<long long> fix_pos;
  DevBuf<int> fix_val;
  auto put = [&](const std::vector<long long>& pos, const std::vector<int>& val, int* dst, int* mk) {
    if (pos.empty()) return;
    fix_pos.upload(pos);
    fix_val.upload(val);
    hipLaunchKernelGGL(k_put, dim3(grid_for((long long)pos.size())), dim3(kThreads), 0, 0, (int)pos.size(),
                       fix_pos.p, fix_val.p, dst, mk);
    HIP_TRY(hipGetLastError());
  };
  {
    const int nn = read_one(scal.p + 1);
    if (nn > kNearCap) {
      err = "too many interpolation entries on cone boundaries for the device cone builder";
      return false;
    }
    std::vector<NearEntry> ne;
    download(ne, near_e, (size_t)nn);
    std::vector<long long> pos(nn);
    std::vector<int> val(nn);
    for (int k = 0; k < nn; k++) {
      const LevelInfo& L = t.lv[ne[k].l];
      const long long q = grid_nonrel(host_grid(ne[k].l), t.box_morton[L.box_off + ne[k].c], in.tx[ne[k].i], in.ty[ne[k].i]);
      if (q < 0) {
        err = "target outside the cone grid of a cousin box";
        return false;
      }
      pos[k] = ne[k].e;
      val[k] = (int)(L.map_off + (long long)ne[k].c * L.nslot + q);
    }
    put(pos, val, ent.p, mark.p);
    t.near_boundary_entries = nn;
  }
  int cone_off = 0;
  for (int l = 2; l <= D; l++) {
    LevelInfo& L = t.lv[l];
    if (l > 2) {
      const LevelInfo& Pl = t.lv[l - 1];
      const long long nq = (long long)Pl.nslot * 4;
      first.alloc(nq);
      hipLaunchKernelGGL(k_fill, dim3(grid_for(nq)), dim3(kThreads), 0, 0, nq, INT_MAX, first.p);
      if (Pl.ncone) {
        hipLaunchKernelGGL(k_proto_first, dim3(grid_for(Pl.ncone)), dim3(kThreads), 0, 0, Pl.ncone, Pl.cone_off,
                           Pl.nrad, Pl.box_off, cone_box.p, cone_posr.p, cone_post.p, box_child.p, first.p);
        HIP_TRY(hipMemset(scal.p + 2, 0, sizeof(int)));
        hipLaunchKernelGGL(k_proto_fill, dim3(grid_for(nq * P)), dim3(kThreads), 0, 0, nq * P, P, t.Ps, l, lv.p,
                           radial.p, box_morton.p, cone_box.p, box_child.p, first.p, xs.p, xa.p,
                           proto.p + L.proto_off, scal.p, near_n.p, scal.p + 2);
        HIP_TRY(hipGetLastError());
        const int nn = read_one(scal.p + 2);
        if (nn > kNearCap) {
          err = "too many protobox nodes on cone boundaries for the device cone builder";
          return false;
        }
        std::vector<NearNode> nd;
        download(nd, near_n, (size_t)nn);
        std::vector<long long> pos(nn);
        std::vector<int> val(nn);
        for (int k = 0; k < nn; k++) {
          const long long qc = nd[k].idx / P, q = qc / 4;
          const int j = (int)(nd[k].idx % P);
          double px, py;
          grid_cheb2cart(host_grid(l - 1), t.box_morton[Pl.box_off + nd[k].pb], q, in.xs[j % t.Ps], in.xa[j / t.Ps],
                         px, py);
          const long long cq = grid_nonrel(host_grid(l), t.box_morton[L.box_off + nd[k].cb], px, py);
          if (cq < 0) {
            err = "parent interpolation node outside the child cone grid";
            return false;
          }
          pos[k] = nd[k].idx;
          val[k] = (int)cq;
        }
        put(pos, val, proto.p + L.proto_off, nullptr);
        t.near_boundary_nodes += nn;
        const long long nm = (long long)Pl.ncone * 4 * P;
        hipLaunchKernelGGL(k_proto_mark, dim3(grid_for(nm)), dim3(kThreads), 0, 0, nm, P, l, lv.p, cone_box.p,
                           cone_posr.p, cone_post.p, box_child.p, proto.p + L.proto_off, mark.p, nprop.p);
      }
      HIP_TRY(hipGetLastError());
      if (read_one(scal.p) & 2) {
        err = "parent interpolation node outside the child cone grid";
        return false;
      }
    }
    const long long n = (long long)L.nbox * L.nslot;
    hipLaunchKernelGGL(k_number_level, dim3(1), dim3(kNumberThreads), 0, 0, L.map_off, n, L.nslot, L.nrad,
                       cone_off, mark.p, dense_map.p, cone_box.p, cone_posr.p, cone_post.p, scal.p + 3 + l);
    HIP_TRY(hipGetLastError());
    const int nc = read_one(scal.p + 3 + l);
    if (nc == 0 && !t.sep) {
      err = "The number of cone segments on any level cannot be 0";
      return false;
    }
    L.cone_off = cone_off;
    L.ncone = nc;
    lvh[l].cone_off = cone_off;
    HIP_TRY(hipMemcpy(lv.p + l, &lvh[l], sizeof(LevelDev), hipMemcpyHostToDevice));
    cone_off += nc;
  }
  t.n_cones = cone_off;
  it_cone.alloc(nent);
  if (nent)
    hipLaunchKernelGGL(k_it_cone, dim3(grid_for(nent)), dim3(kThreads), 0, 0, nent, D, lv.p, ent.p, dense_map.p,
                       it_cone.p);
  HIP_TRY(hipGetLastError());
  download(t.dense_map, dense_map, map_total);
  download(t.proto, proto, proto_total);
  download(t.cone_box, cone_box, cone_off);
  download(t.cone_posr, cone_posr, cone_off);
  download(t.cone_post, cone_post, cone_off);
  download(t.it_cone, it_cone, nent);
  t.n_prop = (long long)read_one(nprop.p);
  t.cone_level.resize(cone_off);
  for (int l = 2; l <= D; l++)
    for (int g = t.lv[l].cone_off; g < t.lv[l].cone_off + t.lv[l].ncone; g++) t.cone_level[g] = (unsigned char)l;
  return true;
}

}  // namespace ifgf

namespace ifgf {

bool tree_groups_device(const GroupJob& job, Tree& t, std::vector<int>& pbox, ConeInput& cin, std::string& err) {
  const int D = t.D;
  const long long nt = job.nt;
  std::vector<GLevel> gh(D + 1);
  for (int l = 2; l <= D; l++) {
    const LevelInfo& L = t.lv[l];
    gh[l] = GLevel{L.min_x, L.min_y, L.bs, job.nbps[l], 0, L.box_off, L.nbox, 0, 0};
  }
  DevBuf<GLevel> glv;
  DevBuf<double> tx, ty;
  DevBuf<long long> box_morton, codes, gcode, it_ptr;
  DevBuf<int> pb, cnt, cz_ptr, cz_list, nb_ptr, nb_list, scal;
  glv.upload(gh);
  tx.upload(std::vector<double>(job.tx, job.tx + nt));
  ty.upload(std::vector<double>(job.ty, job.ty + nt));
  box_morton.upload(t.box_morton);
  codes.alloc((size_t)(D + 1) * nt);
  pb.alloc((size_t)(D + 1) * nt);
  scal.alloc(1 + (size_t)D + 1);  // [0] error flags, [1 + l] groups of field level l
  HIP_TRY(hipMemset(scal.p, 0, scal.n * sizeof(int)));
  hipLaunchKernelGGL(k_target_boxes, dim3(grid_for(nt)), dim3(kThreads), 0, 0, nt, D, job.sep ? 1 : 0, tx.p, ty.p,
                     glv.p, box_morton.p, codes.p, pb.p, scal.p);
  HIP_TRY(hipGetLastError());
  const int fl = read_one(scal.p);
  if (fl & 1) {
    err = "The 2D box index cannot be < 0 or larger than the number of boxes";
    return false;
  }
  if (fl & 2) {
    err = "point outside the relevant boxes of a level";
    return false;
  }
  // groups: the forward map's are its boxes (global box ids), a field's the
  // runs of equal codes of its sorted targets
  long long ngrp = 0;
  if (job.sep) gcode.alloc((size_t)(D + 1) * nt);  // at most nt groups per level
  for (int l = 2; l <= D; l++) {
    gh[l].grp_off = ngrp;
    if (job.sep) {
      hipLaunchKernelGGL(k_group_runs, dim3(1), dim3(kNumberThreads), 0, 0, nt, D, l, codes.p + (size_t)l * nt, pb.p,
                         gcode.p + ngrp, scal.p + 1 + l);
      HIP_TRY(hipGetLastError());
      gh[l].ngrp = read_one(scal.p + 1 + l);
    } else {
      gh[l].ngrp = gh[l].nbox;
    }
    ngrp += gh[l].ngrp;
  }
  glv.upload(gh);
  const long long* grp_code = job.sep ? gcode.p : box_morton.p;
  // cousin lists of every group, level by level
  cnt.alloc(std::max<long long>(ngrp, nt) + 1);
  for (int l = 2; l <= D; l++)
    if (gh[l].ngrp)
      hipLaunchKernelGGL(k_window, dim3(grid_for(gh[l].ngrp)), dim3(kThreads), 0, 0, gh[l].ngrp, l, 1,
                         grp_code + gh[l].grp_off, glv.p, box_morton.p, cnt.p + gh[l].grp_off, (const int*)nullptr,
                         (int*)nullptr);
  DevBuf<long long> tot;
  tot.alloc(ngrp + 1);
  hipLaunchKernelGGL(k_scan<long long>, dim3(1), dim3(kNumberThreads), 0, 0, ngrp, cnt.p, tot.p);
  HIP_TRY(hipGetLastError());
  const long long ncz = read_one(tot.p + ngrp);
  if (ncz > 2147483647LL) {
    err = "too many cousin entries";
    return false;
  }
  cz_ptr.alloc(ngrp + 1);
  hipLaunchKernelGGL(k_scan<int>, dim3(1), dim3(kNumberThreads), 0, 0, ngrp, cnt.p, cz_ptr.p);
  cz_list.alloc(ncz);
  for (int l = 2; l <= D; l++)
    if (gh[l].ngrp)
      hipLaunchKernelGGL(k_window, dim3(grid_for(gh[l].ngrp)), dim3(kThreads), 0, 0, gh[l].ngrp, l, 1,
                         grp_code + gh[l].grp_off, glv.p, box_morton.p, (int*)nullptr, cz_ptr.p + gh[l].grp_off,
                         cz_list.p);
  HIP_TRY(hipGetLastError());
  // leaf neighbourhoods: of every source leaf box, and of a field's leaf groups
  auto neighbourhoods = [&](const long long* gc, int n, std::vector<int>& ptr, std::vector<int>& list) {
    hipLaunchKernelGGL(k_window, dim3(grid_for(n)), dim3(kThreads), 0, 0, n, D, 0, gc, glv.p, box_morton.p, cnt.p,
                       (const int*)nullptr, (int*)nullptr);
    nb_ptr.alloc((size_t)n + 1);
    hipLaunchKernelGGL(k_scan<int>, dim3(1), dim3(kNumberThreads), 0, 0, (long long)n, cnt.p, nb_ptr.p);
    HIP_TRY(hipGetLastError());
    const int nn = read_one(nb_ptr.p + n);
    nb_list.alloc(nn);
    hipLaunchKernelGGL(k_window, dim3(grid_for(n)), dim3(kThreads), 0, 0, n, D, 0, gc, glv.p, box_morton.p,
                       (int*)nullptr, nb_ptr.p, nb_list.p);
    HIP_TRY(hipGetLastError());
    download(ptr, nb_ptr, (size_t)n + 1);
    download(list, nb_list, (size_t)nn);
  };
  neighbourhoods(box_morton.p + gh[D].box_off, gh[D].nbox, t.nb_ptr, t.nb_list);
  if (job.sep) neighbourhoods(gcode.p + gh[D].grp_off, gh[D].ngrp, t.tnb_ptr, t.tnb_list);
  // interpolation entries per target
  if (nt)
    hipLaunchKernelGGL(k_entry_counts, dim3(grid_for(nt)), dim3(kThreads), 0, 0, nt, D, pb.p, glv.p, cz_ptr.p,
                       cnt.p);
  it_ptr.alloc(nt + 1);
  hipLaunchKernelGGL(k_scan<long long>, dim3(1), dim3(kNumberThreads), 0, 0, nt, cnt.p, it_ptr.p);
  HIP_TRY(hipGetLastError());
  download(t.it_ptr, it_ptr, (size_t)nt + 1);
  download(pbox, pb, (size_t)(D + 1) * nt);
  download(cin.cz_ptr, cz_ptr, (size_t)ngrp + 1);
  download(cin.cz_list, cz_list, (size_t)ncz);
  cin.grp_off.assign(D + 2, 0);
  for (int l = 2; l <= D; l++) cin.grp_off[l] = gh[l].grp_off;
  cin.grp_off[D + 1] = ngrp;
  if (job.sep) {
    t.tleaf.resize(nt);
    for (long long i = 0; i < nt; i++) t.tleaf[i] = pbox[(size_t)i * (D + 1) + D];
  }
  return true;
}

}  // namespace ifgf
