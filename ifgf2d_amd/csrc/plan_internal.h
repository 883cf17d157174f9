// Internal definition of the plan object shared by plan.hip and gmres.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ifgf2d.h"
#include "tables.h"
#include "zmath.h"
#include "setup.h"

namespace ifgf {

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(_e) + \
                               " at " #expr);                                     \
    }                                                                              \
  } while (0)


template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t cnt) {
    free();
    n = cnt;
    if (cnt) HIP_TRY(hipMalloc((void**)&p, cnt * sizeof(T)));
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
};


// A second set of the 8-column blocks' per-call buffers, streams and events
// (plans with nrhs_max >= 16): the 8-column blocks of one call alternate
// between the plan's own set on the caller's stream and this one on st, so
// one block's latency-bound level pushes run beside another block's
// bandwidth-bound near field, leaf operator and final sums.
struct BatchLane {
  DevBuf<zd> rho, xt_b, npart_b, fin_pre_b, nearb8, lpart_b, pbuf_b, lcoef8;
  DevBuf<zf> lpart_b32, pbuf_b32, lcoef8_32;
  hipStream_t st = nullptr, aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_start = nullptr, ev_done = nullptr;
  bool ready = false;
  ~BatchLane() {
    for (hipEvent_t e : {ev_fork, ev_join, ev_start, ev_done})
      if (e) (void)hipEventDestroy(e);
    if (aux) (void)hipStreamDestroy(aux);
    if (st) (void)hipStreamDestroy(st);
  }
};

// Device workspaces of the single-system solvers (gmres.hip), kept by the
// plan between calls.
struct SolverCache;
void destroy_solver_cache(SolverCache* c);
// The learned preconditioner's per-curve data (fno.hip), kept by the plan.
struct FnoGeo;
void destroy_fno_geo(FnoGeo* g);

}  // namespace ifgf

struct ifgf_plan {
  using zd = ifgf::zd;
  using zf = ifgf::zf;
  std::string err;
  ifgf_params prm{};
  ifgf::Curve curve;
  ifgf::Geometry geo;
  ifgf::Correction corr;
  ifgf::Tree tree;
  std::vector<int> patch_sorted, excl;
  long long n_near = 0;
  std::vector<ifgf::ConeGeo> cone_geo;
  std::vector<int> cone_q;
  ifgf::ChebTables cheb{};
  bool on_device = false;
  // a field evaluator (ifgf_field_*): the tree holds separate targets; the
  // application has no identity, correction or excluded pairs, and its
  // output is one value per target (tree.M) in the targets' input order
  bool field = false;
  hipStream_t stream = nullptr;
  // stored path: the near field on a second stream beside the leaf operator
  // and the level-push chain (fork after the gather, join before finalize);
  // chosen by the autotuner
  int overlap = 0;  // near field on the aux stream: 0 no, 1 forked after the leaf stage, 2 after the first level push
  int overlap_b = 0;  // the same for the batched column blocks
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // ifgf_apply_graph: one captured application sequence, keyed by its arguments
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void* g_x = nullptr;
  void* g_y = nullptr;
  int g_nrhs = 0, g_flags = 0;
  long long g_ld = 0;
  // device tables
  ifgf::DevBuf<double> px, py, pnx, pny, w;
  ifgf::DevBuf<double> tpx, tpy;  // field: sorted targets
  ifgf::DevBuf<int> patch, perm, inv, leaf_of_point, leaf_begin, nb_ptr, nb_list;
  ifgf::DevBuf<int4> excl_d;
  ifgf::DevBuf<ifgf::ConeGeo> cg;
  ifgf::DevBuf<int> cone_box, cone_qd, box_child, dense_map, proto, it_cone, corr_patch;
  ifgf::DevBuf<long long> it_ptr, corr_ptr;
  ifgf::DevBuf<zd> corr_w;
  ifgf::DevBuf<ifgf::ChebTables> tabs;
  // stored operators (near-field values, folded leaf operator)
  bool stored = false;
  ifgf::DevBuf<long long> nf_goff, lf_off;
  long long near_entries = 0;  // stored near-field values without the group padding
  ifgf::DevBuf<zd> nf_val, lf_val;
  // level-push propagation / interpolation
  ifgf::DevBuf<ifgf::PushTask> push_tasks;
  ifgf::DevBuf<int> push_einfo;      // per push entry: igeo index, or -1 - ptab index
  ifgf::DevBuf<ifgf::Geo4> push_igeo; // interpolation entries' geometry (push order)
  ifgf::DevBuf<ifgf::Geo4> push_ptab; // propagation geometry per (level, q, cp) and node
  ifgf::DevBuf<int> ifac_off;         // per cone: row of ifac_tab (levels < D)
  ifgf::DevBuf<zd> ifac_tab;          // 1 / centre factor per (level, radial interval, s node)
  ifgf::DevBuf<int> push_coff;     // push entries of global cone g: [coff[g], coff[g+1])
  ifgf::DevBuf<int4> push_cidx;    // per (cone, node): pbuf slots of its <= 4 children (0: none)
  ifgf::DevBuf<zd> pbuf;           // slot 0 = 0, then slot 1 + q for push entry q
  long long nslots = 0;            // 1 + push entries
  ifgf::DevBuf<int> it_slot;       // slot of every interpolation entry (it_ptr order)
  std::vector<long long> push_off; // push entries of global cone g: [off[g], off[g+1])
  std::vector<int> push_level_off; // tasks of level index li: [off[li], off[li+1])
  int push_emax = 512;             // entries per push task (autotuned)
  ifgf::DevBuf<ifgf::LeafItem> leaf_items;  // leaf-operator work items
  ifgf::DevBuf<int> leaf_item_off;          // items of leaf cone c: [off[c], off[c+1])
  ifgf::DevBuf<zd> lpart;                   // leaf partial coefficients [item][P]
  ifgf::DevBuf<int4> nf_seg;  // per target: its leaf box's kMaxNb (end, source shift) pairs
  ifgf::DevBuf<ifgf::NearItem> near_items;  // near-field work items
  ifgf::DevBuf<int> near_item_off;          // items of target group g: [off[g], off[g+1])
  ifgf::DevBuf<zd> npart;                   // near-field partial sums [item][8]
  ifgf::DevBuf<zd> fin_pre;                 // split finalize: identity + correction, near, per target
  ifgf::DevBuf<zd> fin_pre_b;               // the same for the batched columns (widest block)
  // several right-hand sides per application: column blocks of rb_max (2, 4
  // or 8; 0 = none) through the batched kernels; push tasks per block width R
  // (index log2 R) with the cone cap of their LDS windows
  int rb_max = 0;
  ifgf::DevBuf<zd> npart_b, lpart_b, pbuf_b, xt_b;
  ifgf::DevBuf<zf> lpart_b32, pbuf_b32;
  ifgf::DevBuf<ifgf::PushTask> btasks[4];
  std::vector<int> blevel_off[4];
  // mixed-precision plans: the FP64 batched blocks' tasks too (their cones per
  // task follow the FP64 LDS windows; IFGF_APPLY_FP64 blocks)
  ifgf::DevBuf<ifgf::PushTask> btasks64[4];
  std::vector<int> blevel_off64[4];
  // the further sets of 8-column buffers: xlane[0] for nrhs_max >= 16,
  // xlane[1] for nrhs_max >= 24 (lane l >= 1 of a call is xlane[l - 1])
  ifgf::BatchLane xlane[2];
  int max_lanes = 3;  // lanes a call's 8-column blocks may use (autotuned: 2 or 3)
  // mixed precision (IFGF_MIXED): FP32 copies of the leaf operator, the
  // propagation / interpolation geometry and inverse node factors, and FP32
  // coefficient workspaces; the FP64 operator stays available (IFGF_APPLY_FP64)
  bool mixed = false;
  ifgf::DevBuf<zf> lf_val32, ifac_tab32, pbuf32, lpart32;
  ifgf::DevBuf<ifgf::Geo4F> push_igeo32, push_ptab32;
  // per-entry push records in push order (no entry-record indirection): the
  // FP32 ones on a mixed plan, the FP64 ones on a plan with batched blocks
  ifgf::DevBuf<ifgf::Geo4F> push_geo32;
  ifgf::DevBuf<ifgf::Geo4> push_geo_b;
  ifgf::DevBuf<ifgf::ChebTablesF> tabs32;
  // 8-column blocks on the matrix cores: near-field and leaf-operator A tiles,
  // their groups and units, the near sums per (target, column), the leaf
  // coefficients per (leaf cone, node, column) and the identity item map the
  // leaf-level push reads them through
  bool mfma8 = false;      // the MFMA tiles are built (nrhs_max >= 8)
  bool use_mfma8 = true;   // 8-column blocks take them (else the VALU batched kernels)
  ifgf::DevBuf<ifgf::MfmaGroup> nm_grp, lm_grp;
  ifgf::DevBuf<ifgf::MfmaUnit> nm_units, lm_units;
  ifgf::DevBuf<int> nm_segs, lm_cone0, lio_id;
  ifgf::DevBuf<double> nm_A, lm_A;
  ifgf::DevBuf<float> lm_A32, nm_A32;  // mixed plans: the tiles rounded to FP32
  ifgf::DevBuf<zd> nearb8, lcoef8;
  ifgf::DevBuf<zf> lcoef8_32;
  int nleaf = 0;
  // lanes per work item of the stored-path kernels
  int near_lanes = 8 /* kNearChunk */, leaf_lanes = 4 /* kLeafChunk */, final_lanes = 8;
  long long stored_bytes = 0;
  // workspaces
  ifgf::DevBuf<zd> rho, sol, coef, xbuf, ybuf;
  // profiling
  // profiling: 7 events per recorded application, read back (and summed) on
  // demand so recording adds no host synchronisation to the applies
  bool profile = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double stage_ms[6] = {0, 0, 0, 0, 0, 0};
  double setup_s[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // plan creation steps (ifgf_plan_setup_times)
  long long prof_count = 0;
  // solvers: cached workspaces, and the last single-system solve's seconds
  // ([0] workspace set-up, [1] host waits on the device, [2] inner cycles, [3] total)
  ifgf::SolverCache* solver = nullptr;
  double solver_s[4] = {0, 0, 0, 0};
  ifgf::FnoGeo* fno_geo = nullptr;
  ~ifgf_plan() {
    ifgf::destroy_solver_cache(solver);
    ifgf::destroy_fno_geo(fno_geo);
    for (auto& e : ev_pool)
      if (e) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (aux) (void)hipStreamDestroy(aux);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace ifgf {


// one device-resident application on stream st (plan.hip); fp64 selects the
// all-FP64 operator on a mixed-precision plan
int apply_device(ifgf_plan* P, const zd* x, zd* y, hipStream_t st, bool fp64 = false);
// nrhs device columns (stride ld): batched blocks of 8 / 4 / 2 where the plan
// has them (plan.hip)
int apply_columns(ifgf_plan* P, const zd* x, long long ld, zd* y, int nrhs, hipStream_t st, bool fp64 = false);
int plan_fail(ifgf_plan* p, int code, const std::string& msg);
// z = M(v): the learned preconditioner of fno.hip on device vectors of a plan's
// curve (patch order), stream-ordered on st
int fno_apply_device(ifgf_fno* F, ifgf_plan* P, const zd* v, zd* z, hipStream_t st);
// The RP correction built on the GPU (precomp.hip); false (err set) when Np
// exceeds the device kernel's limit.
bool build_correction_device(const Geometry& g, const Curve& c, int F, double k, double eta, double p,
                             Correction& out, std::string& err, hipStream_t st);
// The near-singular search on the GPU (precomp.hip), a build_geometry
// NsSearch: the same candidates, test and order as the host loop.
bool near_singular_device(const Curve& c, double delta, Geometry& g, std::string& err);
// Leaf Morton codes and their stable sort on the GPU (morton.hip); false when
// a point lies outside the leaf grid.
bool device_morton_sort(const double* x, const double* y, long long n, int nlevels, std::vector<int>& perm,
                        std::vector<long long>* codes, hipStream_t st);

}  // namespace ifgf
