// C ABI of the MI355X IFGF forward map (include/ifgf2d.h): plan creation,
// device tables, the apply sequence and the auxiliary entry points.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ifgf2d.h"
#include "kernels.h"
#include "kernels_batched.h"
#include "plan_internal.h"
#include "solver.h"

using namespace ifgf;

static thread_local std::string g_last_error;

static int fail(ifgf_plan* p, int code, const std::string& msg) {
  g_last_error = msg;
  if (p) p->err = msg;
  return code;
}
int ifgf::plan_fail(ifgf_plan* p, int code, const std::string& msg) { return fail(p, code, msg); }

static bool build_stored_any(ifgf_plan* P);
static void autotune(ifgf_plan* P);

static void build_device_side(ifgf_plan* P) {
  const Tree& t = P->tree;
  const long long N = t.N;
  const int Pn = t.Ps * t.Pang;
  HIP_TRY(hipStreamCreate(&P->stream));
  HIP_TRY(hipStreamCreate(&P->aux));
  HIP_TRY(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming));
  P->px.upload(t.sx);
  P->py.upload(t.sy);
  P->pnx.upload(t.snx);
  P->pny.upload(t.sny);
  P->w.upload(P->geo.w);
  P->patch.upload(P->patch_sorted);
  P->perm.upload(t.perm);
  P->leaf_begin.upload(t.leaf_begin);
  if (P->field) {
    // targets: inverse sort, leaf groups and their neighbouring source boxes
    P->inv.upload(t.tinv);
    P->leaf_of_point.upload(t.tleaf);
    P->nb_ptr.upload(t.tnb_ptr);
    P->nb_list.upload(t.tnb_list);
    P->tpx.upload(t.tx);
    P->tpy.upload(t.ty);
  } else {
    P->inv.upload(t.inv);
    P->leaf_of_point.upload(t.leaf_of_point);
    P->nb_ptr.upload(t.nb_ptr);
    P->nb_list.upload(t.nb_list);
    std::vector<int4> ex(N);
    for (long long i = 0; i < N; i++)
      ex[i] = make_int4(P->excl[i * 4], P->excl[i * 4 + 1], P->excl[i * 4 + 2], P->excl[i * 4 + 3]);
    P->excl_d.upload(ex);
  }
  P->cg.upload(P->cone_geo);
  P->cone_box.upload(t.cone_box);
  P->cone_qd.upload(P->cone_q);
  P->box_child.upload(t.box_child);
  P->dense_map.upload(t.dense_map);
  P->proto.upload(t.proto);
  P->it_ptr.upload(t.it_ptr);
  P->it_cone.upload(t.it_cone);
  P->corr_ptr.upload(P->corr.ptr);
  P->corr_patch.upload(P->corr.patch);
  P->corr_w.upload(P->corr.w);
  P->tabs.upload(std::vector<ChebTables>{P->cheb});
  const int R = std::max(1, P->prm.nrhs_max);
  P->rho.alloc((size_t)N * R);
  P->sol.alloc((size_t)N * R);
  P->coef.alloc((size_t)t.n_cones * Pn * R);
  P->xbuf.alloc((size_t)N * R);
  P->ybuf.alloc((size_t)t.M * R);
  P->on_device = true;
}

// Per-cone geometry records and the Chebyshev tables (Chebyshev.h:10-68) with
// the transform's alpha factors.
static void build_cone_tables(ifgf_plan* P) {
  const Tree& t = P->tree;
  P->cone_geo.resize(t.n_cones);
  P->cone_q.resize(t.n_cones);
  for (long long g = 0; g < t.n_cones; g++) {
    const LevelInfo& L = t.lv[t.cone_level[g]];
    const int b = t.cone_box[g], posr = t.cone_posr[g], post = t.cone_post[g];
    ConeGeo& G = P->cone_geo[g];
    G.cx = t.box_cx[L.box_off + b];
    G.cy = t.box_cy[L.box_off + b];
    G.h = L.h;
    G.rmin = L.h / t.radial[L.radial_off + posr + 1];
    G.dr = L.h / t.radial[L.radial_off + posr] - G.rmin;
    G.da = M_PI / L.nce;
    G.th0 = G.da * post;
    G.pad = 0;
    P->cone_q[g] = post * L.nrad + posr;
  }
  const int Ps = P->prm.Ps, Pa = P->prm.Pang;
  auto pts = [](int n, double* x) {
    for (int i = 0; i < n; i++) x[n - 1 - i] = std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * n));
  };
  pts(Ps, P->cheb.xs);
  pts(Pa, P->cheb.xa);
  auto tn = [](int n, const double* x, double* T, double scale) {
    std::vector<double> t((size_t)n * n);
    for (int i = 0; i < n; i++) {
      t[i] = 1.0;
      t[n + i] = x[i];
    }
    for (int m = 2; m < n; m++)
      for (int i = 0; i < n; i++) t[m * n + i] = 2.0 * x[i] * t[(m - 1) * n + i] - t[(m - 2) * n + i];
    for (int m = 0; m < n; m++)
      for (int i = 0; i < n; i++) T[m * 16 + i] = t[m * n + i] * (m == 0 ? 1.0 : 2.0) * scale;
  };
  tn(Ps, P->cheb.xs, P->cheb.Ts, 1.0);
  tn(Pa, P->cheb.xa, P->cheb.Ta, 1.0 / (Ps * Pa));
}

extern "C" {

const char* ifgf_last_error(void) { return g_last_error.c_str(); }
const char* ifgf_plan_error(const ifgf_plan* p) { return p ? p->err.c_str() : ""; }

int ifgf_plan_create(const ifgf_surface* s, const ifgf_params* prm, ifgf_plan** out) {
  if (!s || !prm || !out) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  *out = nullptr;
  if (prm->nlevels < 3) return fail(nullptr, IFGF_ERR_LEVELS, "IFGF accelerator requires at least 3 levels.");
  if (prm->Ps < 2 || prm->Ps > 16 || prm->Pang < 2 || prm->Pang > 16)
    return fail(nullptr, IFGF_ERR_ARG, "Ps and Pang must lie in [2, 16]");
  if (prm->formulation < 0 || prm->formulation > 2) return fail(nullptr, IFGF_ERR_ARG, "unknown formulation");
  if (prm->precision != IFGF_FP64 && prm->precision != IFGF_MIXED)
    return fail(nullptr, IFGF_ERR_ARG, "unsupported precision mode");
  if (prm->precision == IFGF_MIXED && (prm->flags & IFGF_PLAN_ON_THE_FLY))
    return fail(nullptr, IFGF_ERR_ARG, "mixed precision needs the stored operators (no IFGF_PLAN_ON_THE_FLY)");
  if (!s->patch_lims || s->npatches < 3 || !(s->delta > 0) || !(s->rp_p >= 2))
    return fail(nullptr, IFGF_ERR_ARG, "invalid surface description");
  std::unique_ptr<ifgf_plan> P(new ifgf_plan);
  P->prm = *prm;
  if (P->prm.nrhs_max < 1) P->prm.nrhs_max = 1;
  std::string err;
  const auto t_start = std::chrono::steady_clock::now();
  auto t_last = t_start;
  auto lap = [&](int step) {
    const auto now = std::chrono::steady_clock::now();
    P->setup_s[step] += std::chrono::duration<double>(now - t_last).count();
    P->setup_s[7] = std::chrono::duration<double>(now - t_start).count();
    t_last = now;
  };
  try {
    if (!make_curve(s->curve_type, s->curve_params, s->ncurve_params, P->curve, err))
      return fail(nullptr, IFGF_ERR_ARG, err);
    const bool dev_geo = (prm->flags & IFGF_PLAN_DEVICE_GEOMETRY) && !(prm->flags & IFGF_PLAN_HOST_ONLY);
    if (!build_geometry(P->curve, s->patch_lims, s->npatches, s->Np, s->delta, P->geo, err,
                        dev_geo ? near_singular_device : nullptr))
      return fail(nullptr, IFGF_ERR_GEOMETRY, err);
    lap(0);
    if ((prm->flags & IFGF_PLAN_DEVICE_PRECOMP) && !(prm->flags & IFGF_PLAN_HOST_ONLY)) {
      if (!build_correction_device(P->geo, P->curve, prm->formulation, prm->k, prm->eta, s->rp_p, P->corr, err,
                                   nullptr))
        return fail(nullptr, IFGF_ERR_ARG, err);
    } else if (!build_correction(P->geo, P->curve, prm->formulation, prm->k, prm->eta, s->rp_p, P->corr, err)) {
      return fail(nullptr, IFGF_ERR_GEOMETRY, err);
    }
    lap(1);
    std::vector<int> dperm;
    if ((prm->flags & IFGF_PLAN_DEVICE_SORT) && !(prm->flags & IFGF_PLAN_HOST_ONLY)) {
      if (prm->nlevels > 16) return fail(nullptr, IFGF_ERR_ARG, "nlevels above 16 (leaf codes are 32-bit)");
      if (!device_morton_sort(P->geo.x.data(), P->geo.y.data(), P->geo.N, prm->nlevels, dperm, nullptr, nullptr))
        return fail(nullptr, IFGF_ERR_GEOMETRY, "The 2D box index cannot be < 0 or larger than the number of boxes");
    }
    const bool dev_tree = (prm->flags & IFGF_PLAN_DEVICE_TREE) && !(prm->flags & IFGF_PLAN_HOST_ONLY);
    if (!build_tree(P->geo.x.data(), P->geo.y.data(), P->geo.nx.data(), P->geo.ny.data(), P->geo.N,
                    prm->nlevels, prm->k, prm->Ps, prm->Pang, P->tree, err, dperm.empty() ? nullptr : &dperm,
                    nullptr, nullptr, 0, dev_tree ? relevant_cones_device : nullptr,
                    dev_tree ? tree_groups_device : nullptr))
      return fail(nullptr, IFGF_ERR_GEOMETRY, err);
    lap(2);
    if (!build_exclusions(P->geo, P->tree, P->patch_sorted, P->excl, P->n_near, err))
      return fail(nullptr, IFGF_ERR_GEOMETRY, err);
    build_cone_tables(P.get());
    lap(3);
    if (!(prm->flags & IFGF_PLAN_HOST_ONLY)) {
      build_device_side(P.get());
      lap(4);
      if (!(prm->flags & IFGF_PLAN_ON_THE_FLY)) {
        if (!build_stored_any(P.get())) return fail(nullptr, IFGF_ERR_ARG, "no kernel instantiation for this (Ps, Pang)");
        lap(5);
        P->overlap = P->overlap_b = (prm->flags & IFGF_PLAN_OVERLAP) ? 2 : 0;
        P->use_mfma8 = !(prm->flags & IFGF_PLAN_NO_MFMA);
        if (prm->flags & IFGF_PLAN_AUTOTUNE) autotune(P.get());
        lap(6);
      }
    }
  } catch (std::exception& e) {
    return fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  *out = P.release();
  return IFGF_OK;
}

void ifgf_plan_destroy(ifgf_plan* p) { delete p; }

long long ifgf_plan_size(const ifgf_plan* p) { return p ? p->geo.N : -1; }

int ifgf_plan_points(const ifgf_plan* p, double* x, double* y, double* nx, double* ny, double* w) {
  if (!p) return fail(nullptr, IFGF_ERR_ARG, "null plan");
  const auto& g = p->geo;
  if (x) std::copy(g.x.begin(), g.x.end(), x);
  if (y) std::copy(g.y.begin(), g.y.end(), y);
  if (nx) std::copy(g.nx.begin(), g.nx.end(), nx);
  if (ny) std::copy(g.ny.begin(), g.ny.end(), ny);
  if (w) std::copy(g.w.begin(), g.w.end(), w);
  return IFGF_OK;
}

int ifgf_plan_stats(const ifgf_plan* p, long long out[20]) {
  if (!p || !out) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  out[0] = p->n_near;
  out[1] = p->tree.n_leaf;
  out[2] = p->tree.n_prop;
  out[3] = p->tree.n_interp;
  out[4] = p->tree.n_cones;
  out[5] = (long long)p->corr.patch.size();
  out[6] = p->tree.nlev;
  out[7] = p->geo.N;
  out[8] = p->tree.lv[p->tree.D].ncone;
  out[9] = p->stored ? 1 : 0;
  out[10] = p->stored ? p->near_entries : 0;
  out[11] = p->stored_bytes;
  out[12] = p->near_lanes;
  out[13] = p->leaf_lanes;
  out[14] = p->push_emax;
  out[15] = p->final_lanes;
  out[16] = p->overlap;
  out[17] = p->mixed ? 1 : 0;
  out[18] = (long long)p->push_ptab.n;
  out[19] = (long long)p->ifac_tab.n;
  return IFGF_OK;
}

int ifgf_plan_sorting(const ifgf_plan* p, long long* sorting) {
  if (!p || !sorting) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  for (long long i = 0; i < p->tree.N; i++) sorting[i] = p->tree.perm[i];
  return IFGF_OK;
}

long long ifgf_plan_level_cones(const ifgf_plan* p, int level, long long* box, long long* nonrel) {
  if (!p || level < 2 || level > p->tree.D) {
    fail(nullptr, IFGF_ERR_ARG, "level out of range");
    return -1;
  }
  const LevelInfo& L = p->tree.lv[level];
  for (int c = 0; c < L.ncone; c++) {
    const int g = L.cone_off + c;
    if (box) box[c] = p->tree.box_morton[L.box_off + p->tree.cone_box[g]];
    if (nonrel) nonrel[c] = (long long)p->tree.cone_post[g] * L.nrad + p->tree.cone_posr[g];
  }
  return L.ncone;
}

long long ifgf_plan_correction(const ifgf_plan* p, long long* ptr, int* patch, double* w) {
  if (!p) {
    fail(nullptr, IFGF_ERR_ARG, "null plan");
    return -1;
  }
  if (ptr) std::copy(p->corr.ptr.begin(), p->corr.ptr.end(), ptr);
  if (patch) std::copy(p->corr.patch.begin(), p->corr.patch.end(), patch);
  if (w) std::memcpy(w, p->corr.w.data(), p->corr.w.size() * sizeof(zd));
  return (long long)p->corr.patch.size();
}

}  // extern "C"

// ------------------------------------------------------------------ apply ---
// lanes that share one target (near field, interpolation) or one leaf cone node
constexpr int kNearLanes = 8;
constexpr int kInterpLanes = 4;
constexpr int kFinalLanes = 8;
template <int P>
struct LeafLanes {
  static constexpr int value = (P * 2 <= 256) ? 2 : 1;
};

// Per-stage timing events of one application: [2s] / [2s+1] bracket stage s
// (gather, near, leaf, propagation, interpolation, finalize) on the stream the
// stage runs on.
static int events_per_apply(const ifgf_plan*) { return 12; }

static hipEvent_t* take_events(ifgf_plan* P) {
  const size_t per = events_per_apply(P);
  if (P->ev_used + per > P->ev_pool.size()) {
    const size_t old = P->ev_pool.size();
    P->ev_pool.resize(old + per * 64, nullptr);
    for (size_t q = old; q < P->ev_pool.size(); q++) HIP_TRY(hipEventCreate(&P->ev_pool[q]));
  }
  hipEvent_t* ev = &P->ev_pool[P->ev_used];
  P->ev_used += per;
  return ev;
}

// paced: the near field beside the level pushes (second stream)
static void launch_near_stored(ifgf_plan* P, hipStream_t st, bool paced = false) {
  const int N = (int)P->tree.M, M = (int)P->near_items.n;
  if (M == 0) return;  // a field whose targets have no source in their leaf neighbourhood
  if (paced)
    hipLaunchKernelGGL(k_near_items<true>, dim3((M + 3) / 4), dim3(256), 0, st, N, M, P->near_items.p, P->nf_seg.p,
                       P->nf_val.p, P->rho.p, P->npart.p);
  else
    hipLaunchKernelGGL(k_near_items<false>, dim3((M + 3) / 4), dim3(256), 0, st, N, M, P->near_items.p, P->nf_seg.p,
                       P->nf_val.p, P->rho.p, P->npart.p);
}

template <int Ps, int Pang, int EPT, typename R>
static void launch_level_push_ept(ifgf_plan* P, hipStream_t st, int li0, int li1) {
  const int D = P->tree.D;
  PushArgsT<R> a{};
  if constexpr (std::is_same<R, float>::value) {
    a = PushArgsT<R>{nullptr, nullptr, P->push_geo32.p, nullptr, P->push_coff.p, P->push_cidx.p,
                     P->leaf_item_off.p, P->lpart32.p, P->tree.lv[D].cone_off, P->ifac_off.p, P->ifac_tab32.p,
                     P->tabs32.p, P->pbuf32.p, 0};
  } else {
    const bool direct = P->push_geo_b.n > 0;
    a = PushArgsT<R>{nullptr, direct ? nullptr : P->push_einfo.p, direct ? P->push_geo_b.p : P->push_igeo.p,
                     direct ? nullptr : P->push_ptab.p, P->push_coff.p, P->push_cidx.p,
                     P->leaf_item_off.p, P->lpart.p, P->tree.lv[D].cone_off, P->ifac_off.p, P->ifac_tab.p,
                     P->tabs.p, P->pbuf.p, 0};
  }
  for (int li = li0; li + 1 < D && li < li1; li++) {
    const int t0 = P->push_level_off[li], nt = P->push_level_off[li + 1] - t0;
    if (nt == 0) continue;
    a.tasks = P->push_tasks.p + t0;
    a.leaf = li == 0;
    hipLaunchKernelGGL((k_level_push<Ps, Pang, EPT, R>), dim3(nt), dim3(256), 0, st, a);
  }
}

// tasks hold at most 256 * EPT entries (push_emax)
template <int Ps, int Pang, typename R>
static void launch_level_push(ifgf_plan* P, hipStream_t st, int li0 = 0, int li1 = 1 << 30) {
  switch (P->push_emax) {
    case 256: launch_level_push_ept<Ps, Pang, 1, R>(P, st, li0, li1); break;
    case 512: launch_level_push_ept<Ps, Pang, 2, R>(P, st, li0, li1); break;
    case 1024: launch_level_push_ept<Ps, Pang, 4, R>(P, st, li0, li1); break;
    default: launch_level_push_ept<Ps, Pang, 8, R>(P, st, li0, li1); break;
  }
}

// One application y = A x on stream st: gather -> near field -> leaf cones ->
// one push launch per level (propagation + interpolation) -> correction,
// identity and unsort. (Schedules that overlapped whole stages on a second
// stream or in shared grids were measured slower on the 64k kite: the
// kernels' times added up under contention.)
// R = double: the all-FP64 operator; R = float: the mixed-precision operator
// (FP32 leaf operator, coefficients, propagation and interpolation).
template <int F, int Ps, int Pang, typename R>
static void launch_ifgf(ifgf_plan* P, const zd* x, zd* y, hipStream_t st) {
  using Z = typename complex_of<R>::type;
  constexpr bool f32 = std::is_same<R, float>::value;
  constexpr int Pn = Ps * Pang;
  constexpr int CPB = 256 / Pn;
  constexpr int kLeafLanes = LeafLanes<Pn>::value;
  const Tree& t = P->tree;
  const int N = (int)t.N, D = t.D;
  const double k = P->prm.k, eta = P->prm.eta;
  const int nb = (N + 255) / 256;
  const bool prof = P->profile;
  hipEvent_t* ev = prof ? take_events(P) : nullptr;
  auto mark = [&](int i) {
    if (prof) HIP_TRY(hipEventRecord(ev[i], st));
  };

  mark(0);
  hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, st, N, x, P->w.p, P->perm.p, P->rho.p);
  mark(1);

  // near field (stored path with overlap: on the second stream, joined
  // before the finalize; not while stage timing is recorded)
  const bool fork = P->stored && P->overlap && !prof;
  const double ident = (F == kSingleLayer) ? 0.0 : 0.5;
  const Z* slots = nullptr;
  if (P->stored) slots = f32 ? (const Z*)P->pbuf32.p : (const Z*)P->pbuf.p;
  const int NT = (int)t.M;
  const zd* fx = P->field ? nullptr : x;
  const long long* fcorr = P->field ? nullptr : P->corr_ptr.p;
  // finalize stage 0 (all), 1 (identity, correction, near field) or 2
  // (interpolation entries), with the plan's lanes per target
  auto finalize = [&](int stage, hipStream_t s) {
#define IFGF_FIN(S)                                                                                          \
  hipLaunchKernelGGL((k_finalize<S, Z>), dim3((NT + 256 / S - 1) / (256 / S)), dim3(256), 0, s, NT, P->geo.Np, \
                     ident, fx, fcorr, P->corr_patch.p, P->corr_w.p, P->inv.p, P->sol.p, P->near_item_off.p,  \
                     P->stored ? P->npart.p : nullptr, P->it_ptr.p, P->it_slot.p, slots, y, stage, P->fin_pre.p)
    if (P->final_lanes == 4) IFGF_FIN(4);
    else if (P->final_lanes == 16) IFGF_FIN(16);
    else IFGF_FIN(8);
#undef IFGF_FIN
  };
  mark(2);
  if (fork) {
    // forked further down
  } else if (P->stored) {
    launch_near_stored(P, st);
  } else {
    hipLaunchKernelGGL((k_near<F, kNearLanes>), dim3((N + 256 / kNearLanes - 1) / (256 / kNearLanes)), dim3(256), 0,
                       st, N, P->px.p, P->py.p, P->pnx.p, P->pny.p, P->patch.p, P->excl_d.p, P->leaf_of_point.p,
                       P->leaf_begin.p, P->nb_ptr.p, P->nb_list.p, P->rho.p, P->sol.p, k, eta);
  }
  mark(3);

  // leaf cones
  mark(4);
  const LevelInfo& LD = t.lv[D];
  if (P->stored) {
    const long long M = P->leaf_items.n;
    const Z* lf = f32 ? (const Z*)P->lf_val32.p : (const Z*)P->lf_val.p;
    Z* lpart = f32 ? (Z*)P->lpart32.p : (Z*)P->lpart.p;
    if (M > 0)
    hipLaunchKernelGGL((k_leaf_items<Pn, Z>), dim3((unsigned)((M * Pn * kLeafChunk + 255) / 256)), dim3(256), 0, st,
                       (int)M, P->leaf_items.p, lf, P->rho.p, lpart);
  } else {
    constexpr int LCPB = 256 / (Pn * kLeafLanes);
    hipLaunchKernelGGL((k_leaf<F, Ps, Pang, kLeafLanes>), dim3((LD.ncone + LCPB - 1) / LCPB), dim3(256), 0, st,
                       LD.ncone, LD.cone_off, P->cg.p, P->cone_box.p, P->leaf_begin.p, P->px.p, P->py.p,
                       P->pnx.p, P->pny.p, P->rho.p, P->coef.p, P->tabs.p, k, eta);
  }
  mark(5);
  // the fork: after the leaf stage (overlap 1) or after the first level push
  // (overlap 2), so that the near-field stream shares the bandwidth with the
  // latency-bound level pushes rather than with the leaf stream
  const int fork_li = P->overlap == 2 ? 1 : 0;
  auto do_fork = [&]() {
    HIP_TRY(hipEventRecord(P->ev_fork, st));
    HIP_TRY(hipStreamWaitEvent(P->aux, P->ev_fork, 0));
    launch_near_stored(P, P->aux, true);
    finalize(1, P->aux);
    HIP_TRY(hipEventRecord(P->ev_join, P->aux));
  };

  // propagation and interpolation: stored path, one push launch per level
  // (both stages in it, timed as the propagation stage); on the fly, the
  // pull chain and one interpolation pass over all points
  mark(6);
  if (P->stored) {
    if (fork) {
      launch_level_push<Ps, Pang, R>(P, st, 0, fork_li);
      do_fork();
      launch_level_push<Ps, Pang, R>(P, st, fork_li);
    } else {
      launch_level_push<Ps, Pang, R>(P, st);
    }
  } else {
    for (int l = D; l >= 3; l--) {
      const LevelInfo& L = t.lv[l];
      const LevelInfo& Pl = t.lv[l - 1];
      hipLaunchKernelGGL((k_prop<Ps, Pang>), dim3((Pl.ncone + CPB - 1) / CPB), dim3(256), 0, st, Pl.ncone,
                         Pl.cone_off, Pl.box_off, Pl.nrad, L.cone_off, L.nslot, P->cg.p, P->cone_box.p,
                         P->cone_qd.p, P->box_child.p, P->proto.p + L.proto_off, P->dense_map.p + L.map_off,
                         P->coef.p, P->tabs.p, k);
    }
  }
  mark(7);
  mark(8);
  if (!P->stored)
    hipLaunchKernelGGL((k_interp<Ps, Pang, kInterpLanes>), dim3((N + 256 / kInterpLanes - 1) / (256 / kInterpLanes)),
                       dim3(256), 0, st, N, P->px.p, P->py.p, P->it_ptr.p, P->it_cone.p, P->cg.p, P->coef.p,
                       P->sol.p, k);
  mark(9);

  // correction, identity and unsort (a field: near + far at its targets);
  // with the near field on the second stream its first part ran there
  if (fork) HIP_TRY(hipStreamWaitEvent(st, P->ev_join, 0));
  mark(10);
  finalize(fork ? 2 : 0, st);
  mark(11);
  HIP_TRY(hipGetLastError());
}

// One application to a block of R columns (x, y column-major, stride ld) with
// the batched kernels; RT = double on a plain plan, float on a mixed plan.
template <int F, int Ps, int Pang, int R, typename RT>
static void launch_batched(ifgf_plan* P, const zd* x, long long ld, zd* y, hipStream_t st, int lane) {
  using Z = typename complex_of<RT>::type;
  constexpr bool f32 = std::is_same<RT, float>::value;
  constexpr int Pn = Ps * Pang;
  constexpr int CB = batched_cones(Pn, R, (int)sizeof(Z));
  constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : 3;
  const Tree& t = P->tree;
  const int N = (int)t.N, D = t.D;
  // the lane's buffers, second stream and events (lane l >= 1: P->xlane[l - 1])
  BatchLane* L2 = lane ? &P->xlane[lane - 1] : nullptr;
  zd* rho = L2 ? L2->rho.p : P->rho.p;
  zd* xt_b = L2 ? L2->xt_b.p : P->xt_b.p;
  zd* npart_b = L2 ? L2->npart_b.p : P->npart_b.p;
  zd* fin_pre_b = L2 ? L2->fin_pre_b.p : P->fin_pre_b.p;
  zd* nearb8 = L2 ? L2->nearb8.p : P->nearb8.p;
  Z* pbuf_lane = f32 ? (Z*)(L2 ? L2->pbuf_b32.p : P->pbuf_b32.p) : (Z*)(L2 ? L2->pbuf_b.p : P->pbuf_b.p);
  Z* lpart_lane = f32 ? (Z*)(L2 ? L2->lpart_b32.p : P->lpart_b32.p) : (Z*)(L2 ? L2->lpart_b.p : P->lpart_b.p);
  Z* lcoef_lane = f32 ? (Z*)(L2 ? L2->lcoef8_32.p : P->lcoef8_32.p) : (Z*)(L2 ? L2->lcoef8.p : P->lcoef8.p);
  hipStream_t aux = L2 ? L2->aux : P->aux;
  hipEvent_t ev_fork = L2 ? L2->ev_fork : P->ev_fork, ev_join = L2 ? L2->ev_join : P->ev_join;
  hipLaunchKernelGGL((k_gather_b<R>), dim3((unsigned)(((long long)N * R + 255) / 256)), dim3(256), 0, st, N, x, ld,
                     P->w.p, P->perm.p, rho, xt_b);
  const int M = (int)P->near_items.n;
  // with overlap: the near field and the identity / correction half of the
  // final sum on the second stream, forked along the push chain as for one column
  const bool fork = P->overlap_b != 0 && M > 0;
  // blocks of 8 columns on the matrix cores (near field and leaf operator)
  const bool mf = R == 8 && P->mfma8 && P->use_mfma8;
  auto near_b = [&](hipStream_t s) {
    if (mf) {
      if constexpr (f32)
        hipLaunchKernelGGL(k_near_mfma<float>, dim3((unsigned)P->nm_units.n), dim3(256), 0, s, (int)P->nm_units.n,
                           P->nm_units.p, P->nm_grp.p, P->nm_segs.p, P->nm_A32.p, rho, nearb8);
      else
        hipLaunchKernelGGL(k_near_mfma<double>, dim3((unsigned)P->nm_units.n), dim3(256), 0, s, (int)P->nm_units.n,
                           P->nm_units.p, P->nm_grp.p, P->nm_segs.p, P->nm_A.p, rho, nearb8);
    } else if (M > 0) {
      hipLaunchKernelGGL((k_near_items_b<R>), dim3((M + 3) / 4), dim3(256), 0, s, N, M, P->near_items.p,
                         P->nf_seg.p, P->nf_val.p, rho, npart_b);
    }
  };
  const double ident = (F == kSingleLayer) ? 0.0 : 0.5;
  const long long nthr = (long long)N * (R / (R < kFinCols ? R : kFinCols));
  Z* pbuf_ = pbuf_lane;
  auto finalize_b = [&](int stage, hipStream_t s) {
#define IFGF_FINB(S)                                                                                              \
  hipLaunchKernelGGL((k_finalize_b<S, R, Z>), dim3((unsigned)((nthr * S + 255) / 256)), dim3(256), 0, s, N,         \
                     P->geo.Np, ident, x, ld, xt_b, P->corr_ptr.p, P->corr_patch.p, P->corr_w.p, P->inv.p,     \
                     P->near_item_off.p, npart_b, P->it_ptr.p, P->it_slot.p, pbuf_, P->rb_max, y, stage,      \
                     fin_pre_b, mf ? nearb8 : nullptr)
    if (P->final_lanes == 4) IFGF_FINB(4);
    else if (P->final_lanes == 16) IFGF_FINB(16);
    else IFGF_FINB(8);
#undef IFGF_FINB
  };
  if (!fork) near_b(st);
  const long long ML = P->leaf_items.n;
  const Z* lf = f32 ? (const Z*)P->lf_val32.p : (const Z*)P->lf_val.p;
  Z* lpart = lpart_lane;
  const int* lio = P->leaf_item_off.p;
  if (mf) {
    // the leaf cones' coefficients for 8 columns, one "item" per cone
    lpart = lcoef_lane;
    lio = P->lio_id.p;
    const RT* A = f32 ? (const RT*)P->lm_A32.p : (const RT*)P->lm_A.p;
    const int nu = (int)P->lm_units.n;
    if (nu > 0)
      hipLaunchKernelGGL((k_leaf_mfma<RT>), dim3((unsigned)((nu + 3) / 4)), dim3(256), 0, st, nu, P->lm_units.p,
                         P->lm_grp.p, A, rho, (Z*)lpart);
  } else {
    hipLaunchKernelGGL((k_leaf_items_b<Pn, R, Z>), dim3((unsigned)((ML * Pn * kLeafChunk + 255) / 256)), dim3(256),
                       0, st, (int)ML, P->leaf_items.p, lf, rho, lpart);
  }
  Z* pbuf = pbuf_lane;
  PushArgsT<RT> a{};
  if constexpr (f32) {
    a = PushArgsT<RT>{nullptr, nullptr, P->push_geo32.p, nullptr, P->push_coff.p, P->push_cidx.p,
                      lio, lpart, t.lv[D].cone_off, P->ifac_off.p, P->ifac_tab32.p, P->tabs32.p, pbuf,
                      0, P->rb_max};
  } else {
    a = PushArgsT<RT>{nullptr, nullptr, P->push_geo_b.p, nullptr, P->push_coff.p, P->push_cidx.p,
                      lio, lpart, t.lv[D].cone_off, P->ifac_off.p, P->ifac_tab.p, P->tabs.p, pbuf, 0,
                      P->rb_max};
  }
  // FP64 blocks on a mixed plan take their own task list
  const bool own64 = !f32 && P->mixed;
  const std::vector<int>& lo = own64 ? P->blevel_off64[LR] : P->blevel_off[LR];
  const int fork_li = P->overlap_b == 2 ? 1 : 0;
  for (int li = 0; li + 1 < D; li++) {
    if (fork && li == fork_li) {
      HIP_TRY(hipEventRecord(ev_fork, st));
      HIP_TRY(hipStreamWaitEvent(aux, ev_fork, 0));
      near_b(aux);
      finalize_b(1, aux);
      HIP_TRY(hipEventRecord(ev_join, aux));
    }
    const int t0 = lo[li], nt = lo[li + 1] - t0;
    if (nt == 0) continue;
    a.tasks = (own64 ? P->btasks64[LR].p : P->btasks[LR].p) + t0;
    a.leaf = li == 0;
    hipLaunchKernelGGL((k_level_push_b<Ps, Pang, R, CB, RT>), dim3(nt), dim3(256), 0, st, a);
  }
  if (fork) HIP_TRY(hipStreamWaitEvent(st, ev_join, 0));
  finalize_b(fork ? 2 : 0, st);
  HIP_TRY(hipGetLastError());
}

template <int F, int Ps, int Pang, typename RT>
static void launch_batched_r(ifgf_plan* P, int R, const zd* x, long long ld, zd* y, hipStream_t st, int lane) {
  if (R == 8) launch_batched<F, Ps, Pang, 8, RT>(P, x, ld, y, st, lane);
  else if (R == 4) launch_batched<F, Ps, Pang, 4, RT>(P, x, ld, y, st, lane);
  else launch_batched<F, Ps, Pang, 2, RT>(P, x, ld, y, st, lane);
}

// Evaluates the stored operators once (plan creation).
// Push tasks per level l = D..2 (level index li = D - l): runs of consecutive
// cones of the level (at most 256 / P cones, at most emax entries); a cone
// with more than emax entries gets tasks of its own, emax entries each.
static void make_push_tasks(const ifgf_plan* P, int emax, int cmax, std::vector<PushTask>& tasks,
                            std::vector<int>& level_off) {
  const Tree& t = P->tree;
  const int D = t.D;
  const std::vector<long long>& off = P->push_off;
  tasks.clear();
  level_off.assign(D, 0);
  for (int li = 0; li + 1 < D; li++) {
    const int l = D - li;
    level_off[li] = (int)tasks.size();
    PushTask cur{0, 0, 0, 0};
    // a task without entries writes nothing and is dropped (the kernels
    // read records at indices clamped into [e0, e0 + ne))
    auto flush = [&] {
      if (cur.nc && cur.ne) tasks.push_back(cur);
      cur = PushTask{0, 0, 0, 0};
    };
    const long long c0 = t.lv[l].cone_off, c1 = c0 + t.lv[l].ncone;
    for (long long c = c0; c < c1; c++) {
      const long long n = off[c + 1] - off[c];
      if (n > emax) {
        flush();
        for (long long s = 0; s < n; s += emax)
          tasks.push_back(PushTask{(int)c, 1, (int)(off[c] + s), (int)std::min<long long>(emax, n - s)});
        continue;
      }
      if (cur.nc == cmax || cur.ne + n > emax) flush();
      if (cur.nc == 0) {
        cur.c0 = (int)c;
        cur.e0 = (int)off[c];
      }
      cur.nc++;
      cur.ne += (int)n;
    }
    flush();
  }
  level_off[D - 1] = (int)tasks.size();
}

static void build_push_tasks(ifgf_plan* P, int emax) {
  if (emax != 256 && emax != 512 && emax != 1024 && emax != 2048) throw std::runtime_error("push task size");
  std::vector<PushTask> tasks;
  make_push_tasks(P, emax, 256 / (P->tree.Ps * P->tree.Pang), tasks, P->push_level_off);
  P->push_tasks.upload(tasks);
  P->push_emax = emax;
}

// Buffers and push tasks of the batched kernels (column blocks up to rb_max).
template <int Ps, int Pang>
static void build_batched(ifgf_plan* P) {
  constexpr int Pn = Ps * Pang;
  const int nr = P->prm.nrhs_max;
  P->rb_max = nr >= 8 ? 8 : nr >= 4 ? 4 : nr >= 2 ? 2 : 0;
  if (!P->rb_max) return;
  const size_t R = (size_t)P->rb_max;
  P->npart_b.alloc(P->npart.n * R);
  P->fin_pre_b.alloc(2 * (size_t)P->tree.M * R);
  P->xt_b.alloc((size_t)P->tree.N * R);
  if (P->mixed) {
    P->lpart_b32.alloc(P->lpart.n * R);
    P->pbuf_b32.alloc((size_t)P->nslots * R);
    HIP_TRY(hipMemsetAsync(P->pbuf_b32.p, 0, P->pbuf_b32.n * sizeof(zf), P->stream));
  }
  // the FP64 blocks (on a mixed plan too: its IFGF_APPLY_FP64 columns, e.g. the
  // refinement's residuals)
  P->lpart_b.alloc(P->lpart.n * R);
  P->pbuf_b.alloc((size_t)P->nslots * R);
  HIP_TRY(hipMemsetAsync(P->pbuf_b.p, 0, P->pbuf_b.n * sizeof(zd), P->stream));
  {
    const long long npush = P->nslots - 1;
    if (P->push_geo_b.n != (size_t)npush) {  // a field with batched blocks
      P->push_geo_b.alloc(npush);
      if (npush > 0)
        hipLaunchKernelGGL(k_geo_direct<double>, dim3((unsigned)((npush + 255) / 256)), dim3(256), 0, P->stream,
                           npush, P->push_einfo.p, P->push_igeo.p, P->push_ptab.p, P->push_geo_b.p);
      HIP_TRY(hipGetLastError());
      P->stored_bytes += (long long)P->push_geo_b.n * (long long)sizeof(Geo4);
    }
  }
  const int zb = P->mixed ? (int)sizeof(zf) : (int)sizeof(zd);
  for (int lr = 1; lr <= 3; lr++) {
    const int r = 1 << lr;
    if (r > P->rb_max) break;
    std::vector<PushTask> tasks;
    make_push_tasks(P, batched_entries(Pn, zb), batched_cones(Pn, r, zb), tasks, P->blevel_off[lr]);
    P->btasks[lr].upload(tasks);
    if (P->mixed) {
      const int z64 = (int)sizeof(zd);
      make_push_tasks(P, batched_entries(Pn, z64), batched_cones(Pn, r, z64), tasks, P->blevel_off64[lr]);
      P->btasks64[lr].upload(tasks);
    }
  }
  P->stored_bytes += (long long)(P->npart_b.n + P->lpart_b.n + P->pbuf_b.n) * (long long)sizeof(zd) +
                     (long long)(P->lpart_b32.n + P->pbuf_b32.n) * (long long)sizeof(zf);
  // leaf-operator A tiles of the 8-column MFMA blocks: one group per leaf box
  // with cones, rows (cone, node) of its cones in order, K over its sources
  if (P->rb_max == 8 && P->nm_units.n > 0) {
    const Tree& t = P->tree;
    const LevelInfo& LD = t.lv[t.D];
    std::vector<MfmaGroup> grp;
    std::vector<MfmaUnit> units;
    std::vector<int> cone0;
    long long aoff = 0;
    for (int c0 = 0; c0 < LD.ncone;) {
      const int b = t.cone_box[LD.cone_off + c0];
      int c1 = c0;
      while (c1 < LD.ncone && t.cone_box[LD.cone_off + c1] == b) c1++;
      const int nsrc = t.leaf_begin[b + 1] - t.leaf_begin[b];
      const int rows = (c1 - c0) * Pn;
      const int g = (int)grp.size();
      grp.push_back(MfmaGroup{aoff, rows, nsrc, c0 * Pn, t.leaf_begin[b]});
      cone0.push_back(c0);
      const int rt = (rows + 15) / 16;
      for (int r = 0; r < rt; r++) units.push_back(MfmaUnit{g, r});
      aoff += (long long)rt * ((nsrc + 1) / 2) * 64;
      c0 = c1;
    }
    P->lm_grp.upload(grp);
    P->lm_units.upload(units);
    P->lm_cone0.upload(cone0);
    std::vector<int> id(LD.ncone + 1);
    for (int c = 0; c <= LD.ncone; c++) id[c] = c;
    P->lio_id.upload(id);
    const unsigned nb = (unsigned)((aoff + 255) / 256);
    if (P->mixed) {
      P->lm_A32.alloc(aoff);
      if (aoff > 0)
        hipLaunchKernelGGL((k_leaf_mfma_build<Pn, zf>), dim3(nb), dim3(256), 0, P->stream, aoff, (int)grp.size(),
                           P->lm_grp.p, P->lm_cone0.p, P->lf_off.p, P->lf_val32.p, P->lm_A32.p);
      P->lcoef8_32.alloc((size_t)LD.ncone * Pn * 8);
      P->stored_bytes += aoff * (long long)sizeof(float);
    }
    P->lm_A.alloc(aoff);
    if (aoff > 0)
      hipLaunchKernelGGL((k_leaf_mfma_build<Pn, zd>), dim3(nb), dim3(256), 0, P->stream, aoff, (int)grp.size(),
                         P->lm_grp.p, P->lm_cone0.p, P->lf_off.p, P->lf_val.p, P->lm_A.p);
    P->lcoef8.alloc((size_t)LD.ncone * Pn * 8);
    P->stored_bytes += aoff * (long long)sizeof(double);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->mfma8 = true;
  }
  // further lanes of 8-column blocks (plans for 16, 24 or more columns)
  for (int xl = 0; xl < 2; xl++) {
    if (P->rb_max != 8 || nr < 16 + 8 * xl) break;
    BatchLane& L = P->xlane[xl];
    L.rho.alloc((size_t)P->tree.N * 8);
    L.xt_b.alloc(P->xt_b.n);
    L.npart_b.alloc(P->npart_b.n);
    L.fin_pre_b.alloc(P->fin_pre_b.n);
    if (P->nearb8.n) L.nearb8.alloc(P->nearb8.n);
    if (P->mixed) {
      L.lpart_b32.alloc(P->lpart_b32.n);
      L.pbuf_b32.alloc(P->pbuf_b32.n);
      HIP_TRY(hipMemsetAsync(L.pbuf_b32.p, 0, L.pbuf_b32.n * sizeof(zf), P->stream));
      if (P->lcoef8_32.n) L.lcoef8_32.alloc(P->lcoef8_32.n);
    }
    L.lpart_b.alloc(P->lpart_b.n);
    L.pbuf_b.alloc(P->pbuf_b.n);
    HIP_TRY(hipMemsetAsync(L.pbuf_b.p, 0, L.pbuf_b.n * sizeof(zd), P->stream));
    if (P->lcoef8.n) L.lcoef8.alloc(P->lcoef8.n);
    if (!L.st) HIP_TRY(hipStreamCreate(&L.st));
    if (!L.aux) HIP_TRY(hipStreamCreate(&L.aux));
    for (hipEvent_t* e : {&L.ev_fork, &L.ev_join, &L.ev_start, &L.ev_done})
      if (!*e) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->stored_bytes += (long long)(L.rho.n + L.xt_b.n + L.npart_b.n + L.fin_pre_b.n + L.nearb8.n + L.lpart_b.n +
                                   L.pbuf_b.n + L.lcoef8.n) * (long long)sizeof(zd) +
                       (long long)(L.lpart_b32.n + L.pbuf_b32.n + L.lcoef8_32.n) * (long long)sizeof(zf);
    L.ready = true;
  }
}

template <int F, int Ps, int Pang>
static void build_stored(ifgf_plan* P) {
  constexpr int Pn = Ps * Pang;
  const Tree& t = P->tree;
  const int D = t.D;
  // N: targets (the sorted sources for the forward map, tree.M separate
  // targets for a field); tleaf / tnb_*: each target's leaf group and the
  // source leaf boxes around it
  const int N = (int)t.M;
  const std::vector<int>& tleaf = P->field ? t.tleaf : t.leaf_of_point;
  const std::vector<int>& tnb_ptr = P->field ? t.tnb_ptr : t.nb_ptr;
  const std::vector<int>& tnb_list = P->field ? t.tnb_list : t.nb_list;
  const double* qx = P->field ? P->tpx.p : P->px.p;
  const double* qy = P->field ? P->tpy.p : P->py.p;
  const double k = P->prm.k, eta = P->prm.eta;
  hipStream_t st = P->stream;
  // near field: segment length of target i = points of its leaf neighbours;
  // groups of kNearGroup targets padded to their longest segment
  std::vector<int> nf_len(N);
  long long near_entries = 0;
  for (int i = 0; i < N; i++) {
    const int b = tleaf[i];
    int c = 0;
    for (int n = tnb_ptr[b]; n < tnb_ptr[b + 1]; n++)
      c += t.leaf_begin[tnb_list[n] + 1] - t.leaf_begin[tnb_list[n]];
    nf_len[i] = c;
    near_entries += c;
  }
  const int ngroups = (N + kNearGroup - 1) / kNearGroup;
  std::vector<long long> nf_goff(ngroups + 1, 0);
  for (int g = 0; g < ngroups; g++) {
    int mx = 0;
    for (int i = g * kNearGroup; i < std::min(N, (g + 1) * kNearGroup); i++) mx = std::max(mx, nf_len[i]);
    const long long chunks = (mx + kNearChunk - 1) / kNearChunk;
    nf_goff[g + 1] = nf_goff[g] + chunks * kNearGroup * kNearChunk;
  }
  P->nf_goff.upload(nf_goff);
  P->near_entries = near_entries;
  {
    std::vector<NearItem> items;
    std::vector<int> item_off(ngroups + 1, 0);
    for (int g = 0; g < ngroups; g++) {
      item_off[g] = (int)items.size();
      const int nchunk = (int)((nf_goff[g + 1] - nf_goff[g]) / (kNearGroup * kNearChunk));
      if (nchunk > 32767) throw std::runtime_error("near-field rows too long");
      for (int u0 = 0; u0 < nchunk; u0 += kNearItemChunks)
        items.push_back(NearItem{nf_goff[g] + (long long)u0 * kNearGroup * kNearChunk, g, (short)u0,
                                 (short)std::min(kNearItemChunks, nchunk - u0)});
    }
    item_off[ngroups] = (int)items.size();
    P->near_items.upload(items);
    P->near_item_off.upload(item_off);
    P->npart.alloc(std::max<size_t>(1, items.size()) * kNearGroup);  // never null: k_finalize tests it
    P->fin_pre.alloc(2 * (size_t)t.M);
  }
  // per target leaf group, its neighbour boxes as (exclusive end in the
  // flattened segment, source index minus flat offset), padded to kMaxNb
  // segments
  P->nleaf = (int)t.leaf_begin.size() - 1;
  const int ntg = (int)tnb_ptr.size() - 1;
  std::vector<int> seg((size_t)ntg * kNbSegInts, 0);
  for (int b = 0; b < ntg; b++) {
    const int nn = tnb_ptr[b + 1] - tnb_ptr[b];
    if (nn > kMaxNb) throw std::runtime_error("a leaf box has more than 9 neighbour boxes");
    int* sg = &seg[(size_t)b * kNbSegInts];
    int cum = 0, q = 0;
    for (int n = tnb_ptr[b]; n < tnb_ptr[b + 1]; n++, q++) {
      const int s0 = t.leaf_begin[tnb_list[n]], len = t.leaf_begin[tnb_list[n] + 1] - s0;
      sg[2 * q + 1] = s0 - cum;
      cum += len;
      sg[2 * q] = cum;
    }
    for (; q < kMaxNb; q++) sg[2 * q] = cum;
  }
  {
    std::vector<int4> per_target((size_t)N * (kNbSegInts / 4));
    for (int i = 0; i < N; i++)
      std::memcpy(&per_target[(size_t)i * (kNbSegInts / 4)], &seg[(size_t)tleaf[i] * kNbSegInts],
                  kNbSegInts * sizeof(int));
    P->nf_seg.upload(per_target);
  }
  P->nf_val.alloc(nf_goff[ngroups]);
  if (nf_goff[ngroups]) HIP_TRY(hipMemsetAsync(P->nf_val.p, 0, nf_goff[ngroups] * sizeof(zd), st));
  hipLaunchKernelGGL((k_near_build<F, kNearLanes>), dim3((N + 256 / kNearLanes - 1) / (256 / kNearLanes)),
                     dim3(256), 0, st, N, qx, qy, P->px.p, P->py.p, P->pnx.p, P->pny.p,
                     P->field ? nullptr : P->patch.p, P->field ? nullptr : P->excl_d.p,
                     P->leaf_of_point.p, P->leaf_begin.p, P->nb_ptr.p, P->nb_list.p, P->nf_goff.p,
                     P->nf_val.p, k, eta);
  HIP_TRY(hipGetLastError());
  // near-field A tiles of the 8-column MFMA blocks: one group per leaf box
  // (its targets are consecutive and share the neighbourhood's sources)
  bool contiguous = true;
  for (int i = 1; i < N && contiguous; i++) contiguous = tleaf[i] >= tleaf[i - 1];
  if (P->prm.nrhs_max >= 8 && !P->field && contiguous && N > 0) {
    std::vector<MfmaGroup> grp;
    std::vector<MfmaUnit> units;
    long long aoff = 0;
    for (int i0 = 0; i0 < N;) {
      int i1 = i0;
      while (i1 < N && tleaf[i1] == tleaf[i0]) i1++;
      const int b = tleaf[i0];
      const int nsrc = seg[(size_t)b * kNbSegInts + 2 * (kMaxNb - 1)];
      const int g = (int)grp.size();
      grp.push_back(MfmaGroup{aoff, i1 - i0, nsrc, i0, b});
      const int rt = (i1 - i0 + 15) / 16;
      for (int r = 0; r < rt; r++) units.push_back(MfmaUnit{g, r});
      aoff += (long long)rt * ((nsrc + 1) / 2) * 64;
      i0 = i1;
    }
    P->nm_grp.upload(grp);
    P->nm_units.upload(units);
    P->nm_segs.upload(seg);
    P->nm_A.alloc(aoff);
    if (aoff > 0)
      hipLaunchKernelGGL(k_near_mfma_build, dim3((unsigned)((aoff + 255) / 256)), dim3(256), 0, st, aoff,
                         (int)grp.size(), P->nm_grp.p, P->nf_goff.p, P->nf_val.p, P->nm_A.p);
    HIP_TRY(hipGetLastError());
    P->nearb8.alloc((size_t)N * 8);
  }
  // leaf operator: P x n_src per leaf cone
  const LevelInfo& LD = t.lv[D];
  // (sources padded to a multiple of kLeafChunk; lf_src = (first source, count))
  std::vector<long long> lf_off(LD.ncone + 1, 0);
  std::vector<int2> lf_src(LD.ncone);
  for (int c = 0; c < LD.ncone; c++) {
    const int b = t.cone_box[LD.cone_off + c];
    const int ns = t.leaf_begin[b + 1] - t.leaf_begin[b];
    lf_src[c] = make_int2(t.leaf_begin[b], ns);
    lf_off[c + 1] = lf_off[c] + (long long)Pn * ((ns + kLeafChunk - 1) / kLeafChunk * kLeafChunk);
  }
  P->lf_off.upload(lf_off);
  {
    std::vector<LeafItem> items;
    std::vector<int> item_off(LD.ncone + 1, 0);
    for (int c = 0; c < LD.ncone; c++) {
      item_off[c] = (int)items.size();
      for (int s = 0; s < lf_src[c].y; s += kLeafItemTiles * kLeafChunk)
        items.push_back(LeafItem{lf_off[c] + (long long)s * Pn, lf_src[c].x + s,
                                 std::min(kLeafItemTiles * kLeafChunk, lf_src[c].y - s)});
    }
    item_off[LD.ncone] = (int)items.size();
    P->leaf_items.upload(items);
    P->leaf_item_off.upload(item_off);
    P->lpart.alloc(items.size() * Pn);
  }
  P->lf_val.alloc(lf_off[LD.ncone]);
  {
    DevBuf<zd> tmp;
    tmp.alloc(lf_off[LD.ncone]);
    const long long nitems = lf_off[LD.ncone];
    if (nitems > 0) {
    hipLaunchKernelGGL((k_leaf_build_values<F, Ps, Pang>), dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0,
                       st, LD.ncone, LD.cone_off, P->cg.p, P->cone_box.p, P->leaf_begin.p, P->lf_off.p,
                       P->lf_off.p, P->px.p, P->py.p, P->pnx.p, P->pny.p, P->tabs.p, tmp.p, nitems, k, eta);
    hipLaunchKernelGGL((k_leaf_build_transform<Ps, Pang>), dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0,
                       st, nitems / Pn, tmp.p, P->tabs.p, P->lf_val.p);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
  }
  // interpolation geometry per (target, cousin cone), in it_ptr order
  DevBuf<Geo4> it_geo;
  it_geo.alloc(t.it_ptr[N]);
  hipLaunchKernelGGL((k_interp_build<kInterpLanes>), dim3((N + 256 / kInterpLanes - 1) / (256 / kInterpLanes)),
                     dim3(256), 0, st, N, qx, qy, P->it_ptr.p, P->it_cone.p, P->cg.p, it_geo.p, k);
  HIP_TRY(hipGetLastError());
  // propagation geometry per (parent node, relevant child): entry
  // pr_ptr[g] + k P + j holds node j of parent cone g in its k-th child
  std::vector<long long> pr_ptr(t.n_cones + 1, 0);
  for (long long g = 0; g < t.n_cones; g++) {
    long long c = 0;
    const int l = t.cone_level[g];
    if (l < D) {
      const int pb = t.cone_box[g];
      for (int cp = 0; cp < 4; cp++) c += t.box_child[(size_t)(t.lv[l].box_off + pb) * 4 + cp] >= 0 ? Pn : 0;
    }
    pr_ptr[g + 1] = pr_ptr[g] + c;
  }
  const long long npr = pr_ptr[t.n_cones], nit = t.it_ptr[N];
  DevBuf<long long> d_pr_ptr;
  DevBuf<int> d_pr_cone;
  DevBuf<Geo4> pr_geo;
  d_pr_ptr.upload(pr_ptr);
  d_pr_cone.alloc(npr);
  pr_geo.alloc(npr);
  DevBuf<zd> pn_ifac;
  pn_ifac.alloc((size_t)t.n_cones * Pn);
  for (int l = D; l >= 3; l--) {
    const LevelInfo& L = t.lv[l];
    const LevelInfo& Pl = t.lv[l - 1];
    if (Pl.ncone == 0) continue;  // a field level without cones
    hipLaunchKernelGGL((k_prop_build<Ps, Pang>), dim3((Pl.ncone * Pn + 255) / 256), dim3(256), 0, st, Pl.ncone,
                       Pl.cone_off, Pl.box_off, L.cone_off, L.nslot, P->cg.p, P->cone_box.p, P->cone_qd.p,
                       P->box_child.p, P->proto.p + L.proto_off, P->dense_map.p + L.map_off, P->tabs.p,
                       d_pr_ptr.p, d_pr_cone.p, pr_geo.p, pn_ifac.p, k);
  }
  HIP_TRY(hipGetLastError());
  std::vector<int> pr_cone(npr);
  if (npr) HIP_TRY(hipMemcpyAsync(pr_cone.data(), d_pr_cone.p, npr * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // push entries grouped by the cone that evaluates them: first the parent
  // nodes that fall in the cone, then the (target, cousin) interpolation
  // entries, then one padding entry if the cone's run is odd, so that every
  // run has even length and the push kernels' pairs of consecutive entries
  // never span two cones (a padding entry is evaluated with a valid record
  // and writes a slot nobody reads). Push entry q writes slot 1 + q of pbuf
  // (slot 0 stays zero), so a task's writes are one contiguous run; a parent
  // node reads its k-th child's value through cidx[node][k] (0 for an absent
  // child) and the finalize a target's interpolations through it_slot.
  if (npr + nit + t.n_cones + 1 > 2147483647LL)
    throw std::runtime_error("too many propagation and interpolation entries");
  const long long NC = t.n_cones;
  // Stable counting sort of the push entries by the cone that evaluates them:
  // a cone's propagation entries in entry order, then its interpolation
  // entries in entry order. Chunks of entries per thread: per-chunk counts,
  // their prefix over (cone, chunk), then every chunk scatters its own.
  const int nth = std::max(1, std::min(omp_get_max_threads(), 64));
  std::vector<int> qprop(npr), qint(nit);  // push position of every entry
  std::vector<long long> off(NC + 1, 0), npro(NC, 0), nint(NC, 0);
  {
    std::vector<std::vector<int>> hp(nth, std::vector<int>(NC, 0)), hi(nth, std::vector<int>(NC, 0));
    auto chunk = [](long long n, int T, int c, long long& b0, long long& b1) {
      b0 = n * c / T;
      b1 = n * (c + 1) / T;
    };
    // chunk c is a loop index, not a thread number: every chunk is counted
    // and scattered whatever team size the runtime grants
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int c = 0; c < nth; c++) {
      long long b0, b1;
      chunk(npr, nth, c, b0, b1);
      for (long long e = b0; e < b1; e++) hp[c][pr_cone[e]]++;
      chunk(nit, nth, c, b0, b1);
      for (long long e = b0; e < b1; e++) hi[c][t.it_cone[e]]++;
    }
    for (long long g = 0; g < NC; g++) {
      long long sp = 0, si = 0;
      for (int c = 0; c < nth; c++) {
        sp += hp[c][g];
        si += hi[c][g];
      }
      npro[g] = sp;
      nint[g] = si;
      off[g + 1] = off[g] + sp + si + ((sp + si) & 1);
    }
    // hp / hi become each chunk's first position in the cone's run
#pragma omp parallel for schedule(static) num_threads(nth)
    for (long long g = 0; g < NC; g++) {
      long long rp = off[g], ri = off[g] + npro[g];
      for (int c = 0; c < nth; c++) {
        const long long cp = hp[c][g], ci = hi[c][g];
        hp[c][g] = (int)rp;
        hi[c][g] = (int)ri;
        rp += cp;
        ri += ci;
      }
    }
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int c = 0; c < nth; c++) {
      long long b0, b1;
      chunk(npr, nth, c, b0, b1);
      for (long long e = b0; e < b1; e++) qprop[e] = hp[c][pr_cone[e]]++;
      chunk(nit, nth, c, b0, b1);
      for (long long e = b0; e < b1; e++) qint[e] = hi[c][t.it_cone[e]]++;
    }
  }
  const long long npush = off[NC];
  // padding entries read the first interpolation record (or propagation row)
  std::vector<int> einfo(npush, nit > 0 ? 0 : -1);
  std::vector<int4> cidx((size_t)NC * Pn, make_int4(0, 0, 0, 0));
  // propagation geometry shared per (level, parent non-relevant id q, child
  // position cp): pair ids in order of first use, each gathered from its
  // first parent cone's entries
  // (pairs of parent level l: [pair_base[l], pair_base[l + 1]), key q * 4 + cp)
  std::vector<long long> pair_base(D + 1, 0);
  for (int l = 2; l < D; l++) pair_base[l + 1] = pair_base[l] + (long long)t.lv[l].nslot * 4;
  std::vector<int> pair_id(pair_base[D], -1);
  std::vector<long long> ptab_src;
  std::vector<int> gid((size_t)NC * 4, -1);  // pair id of parent cone g's k-th child
  for (long long g = 0; g < NC; g++) {
    const int nch = (int)((pr_ptr[g + 1] - pr_ptr[g]) / Pn);
    const int l = t.cone_level[g];
    int kc = 0;
    for (int cp = 0; cp < 4 && kc < nch; cp++) {
      if (t.box_child[(size_t)(t.lv[l].box_off + t.cone_box[g]) * 4 + cp] < 0) continue;
      int& id = pair_id[pair_base[l] + (long long)P->cone_q[g] * 4 + cp];
      if (id < 0) {
        id = (int)(ptab_src.size() / Pn);
        for (int j = 0; j < Pn; j++) ptab_src.push_back(pr_ptr[g] + (long long)kc * Pn + j);
      }
      gid[(size_t)g * 4 + kc] = id;
      kc++;
    }
  }
#pragma omp parallel for schedule(static) num_threads(nth)
  for (long long g = 0; g < NC; g++) {
    const int nch = (int)((pr_ptr[g + 1] - pr_ptr[g]) / Pn);
    for (int kc = 0; kc < nch; kc++) {
      const int id = gid[(size_t)g * 4 + kc];
      for (int j = 0; j < Pn; j++) {
        const long long e = pr_ptr[g] + (long long)kc * Pn + j;
        const int q = qprop[e];
        einfo[q] = -1 - (id * Pn + j);
        int* c = &cidx[(size_t)g * Pn + j].x;
        c[kc] = 1 + q;
      }
    }
  }
  // interpolation entries: slot of each, and their geometry in push order (a
  // cone's interpolation entries follow its propagation entries, so an
  // entry's rank among all interpolation slots is its cone's base plus its
  // place in the cone's run)
  std::vector<int> it_slot(nit);
  std::vector<long long> ibase(NC + 1, 0);
  for (long long g = 0; g < NC; g++) ibase[g + 1] = ibase[g] + nint[g];
  std::vector<long long> igeo_src(nit);
#pragma omp parallel for schedule(static) num_threads(nth)
  for (long long e = 0; e < nit; e++) {
    const int q = qint[e];
    const int g = t.it_cone[e];
    const long long r = ibase[g] + (q - off[g] - npro[g]);
    it_slot[e] = 1 + q;
    einfo[q] = (int)r;
    igeo_src[r] = e;
  }
  // inverse node factors per (level, radial interval, s node), gathered from
  // the first parent cone of each (level, radial interval)
  std::vector<int> ifac_off(t.n_cones, 0);
  std::vector<long long> ifac_src;
  {
    std::vector<long long> rbase(D + 1, 0);
    for (int l = 2; l < D; l++) rbase[l + 1] = rbase[l] + t.lv[l].nrad;
    std::vector<int> row(rbase[D], -1);
    for (long long g = 0; g < t.n_cones; g++) {
      const int l = t.cone_level[g];
      if (l >= D) continue;
      int& r = row[rbase[l] + t.cone_posr[g]];
      if (r < 0) {
        r = (int)ifac_src.size();
        for (int js = 0; js < Ps; js++) ifac_src.push_back(g * Pn + js);
      }
      ifac_off[g] = r;
    }
  }
  P->it_slot.upload(it_slot);
  P->push_cidx.upload(cidx);
  P->push_off = off;
  {
    std::vector<int> coff(off.begin(), off.end());
    P->push_coff.upload(coff);
  }
  P->push_einfo.upload(einfo);
  P->ifac_off.upload(ifac_off);
  {
    auto take = [&](auto* in, auto& out, const std::vector<long long>& idx) {
      out.alloc(idx.size());
      if (idx.empty()) return;
      DevBuf<long long> d_idx;
      d_idx.upload(idx);
      using T = std::remove_pointer_t<decltype(in)>;
      hipLaunchKernelGGL(k_take<T>, dim3((unsigned)((idx.size() + 255) / 256)), dim3(256), 0, st,
                         (long long)idx.size(), d_idx.p, (const T*)in, out.p);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(st));
    };
    take(it_geo.p, P->push_igeo, igeo_src);
    take(pr_geo.p, P->push_ptab, ptab_src);
    take(pn_ifac.p, P->ifac_tab, ifac_src);
  }
  P->nslots = 1 + npush;
  P->pbuf.alloc(P->nslots);
  HIP_TRY(hipMemset(P->pbuf.p, 0, P->nslots * sizeof(zd)));
  build_push_tasks(P, P->push_emax);
  P->stored_bytes = (long long)(nf_goff[ngroups] + lf_off[LD.ncone] + P->ifac_tab.n + P->nslots) *
                        (long long)sizeof(zd) +
                    (long long)(P->push_igeo.n + P->push_ptab.n) * (long long)sizeof(Geo4) +
                    (npush + nit + t.n_cones) * (long long)sizeof(int) +
                    (long long)t.n_cones * Pn * (long long)sizeof(int4);
  P->stored_bytes += (long long)P->nm_A.n * (long long)sizeof(double);  // the 8-column near tiles
  if (P->prm.precision == IFGF_MIXED) {
    // FP32 copies: leaf operator, propagation / interpolation geometry
    // (factors evaluated in FP64 above, rounded here), inverse node factors
    auto round = [&](const void* in, void* out, long long npairs) {
      if (npairs)
        hipLaunchKernelGGL(k_round_f32, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, st, npairs,
                           (const dbl2*)in, (flt2*)out);
    };
    P->lf_val32.alloc(P->lf_val.n);
    round(P->lf_val.p, P->lf_val32.p, (long long)P->lf_val.n);
    P->push_igeo32.alloc(P->push_igeo.n);
    round(P->push_igeo.p, P->push_igeo32.p, 2 * (long long)P->push_igeo.n);
    P->push_ptab32.alloc(P->push_ptab.n);
    round(P->push_ptab.p, P->push_ptab32.p, 2 * (long long)P->push_ptab.n);
    P->push_geo32.alloc(npush);
    if (npush > 0)
      hipLaunchKernelGGL(k_geo_direct<float>, dim3((unsigned)((npush + 255) / 256)), dim3(256), 0, st, npush,
                         P->push_einfo.p, P->push_igeo32.p, P->push_ptab32.p, P->push_geo32.p);
    P->ifac_tab32.alloc(P->ifac_tab.n);
    round(P->ifac_tab.p, P->ifac_tab32.p, (long long)P->ifac_tab.n);
    if (P->nm_A.n) {  // the 8-column near tiles (a multiple of 64 values)
      P->nm_A32.alloc(P->nm_A.n);
      round(P->nm_A.p, P->nm_A32.p, (long long)P->nm_A.n / 2);
      P->stored_bytes += (long long)P->nm_A32.n * (long long)sizeof(float);
    }
    P->pbuf32.alloc(P->pbuf.n);
    HIP_TRY(hipMemsetAsync(P->pbuf32.p, 0, P->pbuf32.n * sizeof(zf), st));
    P->lpart32.alloc(P->lpart.n);
    ChebTablesF tf;
    for (int q = 0; q < 16; q++) {
      tf.xs[q] = (float)P->cheb.xs[q];
      tf.xa[q] = (float)P->cheb.xa[q];
    }
    for (int q = 0; q < 256; q++) {
      tf.Ts[q] = (float)P->cheb.Ts[q];
      tf.Ta[q] = (float)P->cheb.Ta[q];
    }
    P->tabs32.upload(std::vector<ChebTablesF>{tf});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    P->stored_bytes += (long long)(P->lf_val32.n + P->ifac_tab32.n + P->pbuf32.n) * (long long)sizeof(zf) +
                       (long long)(P->push_igeo32.n + P->push_ptab32.n + P->push_geo32.n) * (long long)sizeof(Geo4F);
    P->mixed = true;
  }
  // Forward-map plans read per-entry FP64 records in push order (no
  // entry-record indirection in front of the geometry: 3 us per application
  // faster at the 64k kite); a field keeps the shared tables, whose
  // propagation part is most of its entries.
  if (!P->field) {
    P->push_geo_b.alloc(npush);
    if (npush > 0)
      hipLaunchKernelGGL(k_geo_direct<double>, dim3((unsigned)((npush + 255) / 256)), dim3(256), 0, st, npush,
                         P->push_einfo.p, P->push_igeo.p, P->push_ptab.p, P->push_geo_b.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    P->stored_bytes += (long long)npush * (long long)sizeof(Geo4) -
                       (long long)(P->push_igeo.n + P->push_ptab.n) * (long long)sizeof(Geo4) -
                       npush * (long long)sizeof(int);
    P->push_einfo.free();
    P->push_igeo.free();
    P->push_ptab.free();
    if (P->mixed) {
      P->stored_bytes -= (long long)(P->push_igeo32.n + P->push_ptab32.n) * (long long)sizeof(Geo4F);
      P->push_igeo32.free();
      P->push_ptab32.free();
    }
  }
  build_batched<Ps, Pang>(P);
  HIP_TRY(hipStreamSynchronize(st));
  P->stored = true;
}

template <int F>
static bool dispatch_build(ifgf_plan* P) {
  const int a = P->prm.Ps, b = P->prm.Pang;
#define IFGF_CASE(A, B)            \
  if (a == A && b == B) {          \
    build_stored<F, A, B>(P);      \
    return true;                   \
  }
  IFGF_CASE(5, 5)
  IFGF_CASE(3, 5)
  IFGF_CASE(5, 7)
  IFGF_CASE(7, 7)
  IFGF_CASE(9, 9)
  IFGF_CASE(11, 11)
#undef IFGF_CASE
  return false;
}

static bool build_stored_any(ifgf_plan* P) {
  switch (P->prm.formulation) {
    case kSingleLayer: return dispatch_build<kSingleLayer>(P);
    case kDoubleLayer: return dispatch_build<kDoubleLayer>(P);
    default: return dispatch_build<kCombined>(P);
  }
}

template <int F>
static bool dispatch_p(ifgf_plan* P, const zd* x, zd* y, hipStream_t st, bool fp64) {
  const int a = P->prm.Ps, b = P->prm.Pang;
  const bool f32 = P->mixed && !fp64;
#define IFGF_CASE(A, B)                                   \
  if (a == A && b == B) {                                 \
    if (f32) launch_ifgf<F, A, B, float>(P, x, y, st);    \
    else launch_ifgf<F, A, B, double>(P, x, y, st);       \
    return true;                                          \
  }
  IFGF_CASE(5, 5)
  IFGF_CASE(3, 5)
  IFGF_CASE(5, 7)
  IFGF_CASE(7, 7)
  IFGF_CASE(9, 9)
  IFGF_CASE(11, 11)
#undef IFGF_CASE
  return false;
}

template <int F>
static bool dispatch_b(ifgf_plan* P, int R, const zd* x, long long ld, zd* y, hipStream_t st, int lane, bool fp64) {
  const int a = P->prm.Ps, b = P->prm.Pang;
#define IFGF_CASE(A, B)                                              \
  if (a == A && b == B) {                                            \
    if (P->mixed && !fp64) launch_batched_r<F, A, B, float>(P, R, x, ld, y, st, lane);  \
    else launch_batched_r<F, A, B, double>(P, R, x, ld, y, st, lane);      \
    return true;                                                     \
  }
  IFGF_CASE(5, 5)
  IFGF_CASE(3, 5)
  IFGF_CASE(5, 7)
  IFGF_CASE(7, 7)
  IFGF_CASE(9, 9)
  IFGF_CASE(11, 11)
#undef IFGF_CASE
  return false;
}

// one block of R (2, 4 or 8) columns through the batched kernels
static bool dispatch_block(ifgf_plan* P, int R, const zd* x, long long ld, zd* y, hipStream_t st, int lane = 0,
                           bool fp64 = false) {
  switch (P->prm.formulation) {
    case kSingleLayer: return dispatch_b<kSingleLayer>(P, R, x, ld, y, st, lane, fp64);
    case kDoubleLayer: return dispatch_b<kDoubleLayer>(P, R, x, ld, y, st, lane, fp64);
    default: return dispatch_b<kCombined>(P, R, x, ld, y, st, lane, fp64);
  }
}

static bool dispatch(ifgf_plan* P, const zd* x, zd* y, hipStream_t st, bool fp64 = false) {
  switch (P->prm.formulation) {
    case kSingleLayer: return dispatch_p<kSingleLayer>(P, x, y, st, fp64);
    case kDoubleLayer: return dispatch_p<kDoubleLayer>(P, x, y, st, fp64);
    default: return dispatch_p<kCombined>(P, x, y, st, fp64);
  }
}

// accumulate the recorded applications' stage times and recycle their events
// Times whole applications on the plan's stream for each candidate lane count
// of each stored-path stage (the others held fixed) and keeps the fastest.
static void autotune(ifgf_plan* P) {
  hipStream_t st = P->stream;
  const long long N = P->geo.N;
  std::vector<zd> xh(N);
  for (long long i = 0; i < N; i++) xh[i] = zd{std::cos(0.37 * i), std::sin(0.11 * i)};
  HIP_TRY(hipMemcpy(P->xbuf.p, xh.data(), N * sizeof(zd), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  // the fastest of three rounds of ten applications, after two warm-up ones
  auto time_apply = [&]() {
    for (int w = 0; w < 2; w++) dispatch(P, P->xbuf.p, P->ybuf.p, st);
    float best = 1e30f;
    for (int round = 0; round < 3; round++) {
      HIP_TRY(hipEventRecord(e0, st));
      for (int r = 0; r < 10; r++) dispatch(P, P->xbuf.p, P->ybuf.p, st);
      HIP_TRY(hipEventRecord(e1, st));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
    return best;
  };
  // Two interleaved sweeps over a knob's candidates, the faster time of each
  // kept; starting from the knob's default, a candidate (in list order)
  // replaces the current choice only when it is more than 1% faster than
  // it (near-ties otherwise flip with timing noise between runs).
  // (timer: one-column applications unless given)
  auto pick = [&](int def, const std::vector<int>& cand, const std::function<void(int)>& set,
                  const std::function<float()>& timer = {}) {
    std::vector<float> tm(cand.size(), 1e30f);
    for (int sweep = 0; sweep < 2; sweep++)
      for (size_t q = 0; q < cand.size(); q++) {
        set(cand[q]);
        tm[q] = std::min(tm[q], timer ? timer() : time_apply());
      }
    float tdef = 1e30f;
    for (size_t q = 0; q < cand.size(); q++)
      if (cand[q] == def) tdef = tm[q];
    int best = def;
    float tb = tdef;
    for (size_t q = 0; q < cand.size(); q++)
      if (cand[q] != def && tm[q] < 0.99f * tb) {
        tb = tm[q];
        best = cand[q];
      }
    set(best);
  };
  // entries per level-push task: fewer give more, shorter blocks
  if (P->stored) pick(P->push_emax, {256, 512, 1024, 2048}, [&](int c) { build_push_tasks(P, c); });
  pick(P->final_lanes, {4, 8, 16}, [&](int c) { P->final_lanes = c; });
  pick(P->overlap, {0, 1, 2}, [&](int c) { P->overlap = c; });
  // the batched blocks' second stream, timed on one block of the widest width
  P->overlap_b = P->overlap;
  if (P->stored && P->rb_max > 0) {
    const int R = P->rb_max;
    DevBuf<zd> xb, yb;
    xb.alloc((size_t)N * R);
    yb.alloc((size_t)N * R);
    for (int c = 0; c < R; c++)
      HIP_TRY(hipMemcpy(xb.p + (size_t)c * N, xh.data(), N * sizeof(zd), hipMemcpyHostToDevice));
    auto time_block = [&]() {
      for (int w = 0; w < 2; w++) dispatch_block(P, R, xb.p, N, yb.p, st);
      float best = 1e30f;
      for (int round = 0; round < 3; round++) {
        HIP_TRY(hipEventRecord(e0, st));
        for (int r = 0; r < 3; r++) dispatch_block(P, R, xb.p, N, yb.p, st);
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
      }
      return best;
    };
    int best = P->overlap_b;
    float tbest = 1e30f;
    for (int c : {0, 1, 2}) {
      P->overlap_b = c;
      const float t = time_block();
      if (t < tbest) {
        tbest = t;
        best = c;
      }
    }
    P->overlap_b = best;
    HIP_TRY(hipStreamSynchronize(st));
    // a third lane, timed on a call of 24 columns (three 8-column blocks)
    if (P->xlane[1].ready) {
      DevBuf<zd> x3, y3;
      x3.alloc((size_t)N * 24);
      y3.alloc((size_t)N * 24);
      for (int c = 0; c < 24; c++)
        HIP_TRY(hipMemcpy(x3.p + (size_t)c * N, xh.data(), N * sizeof(zd), hipMemcpyHostToDevice));
      pick(P->max_lanes, {2, 3}, [&](int c) { P->max_lanes = c; }, [&]() {
        float best3 = 1e30f;
        apply_columns(P, x3.p, N, y3.p, 24, st, false);
        for (int round = 0; round < 2; round++) {
          HIP_TRY(hipEventRecord(e0, st));
          for (int r = 0; r < 2; r++) apply_columns(P, x3.p, N, y3.p, 24, st, false);
          HIP_TRY(hipEventRecord(e1, st));
          HIP_TRY(hipEventSynchronize(e1));
          float ms = 0.f;
          HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
          best3 = std::min(best3, ms);
        }
        return best3;
      });
      HIP_TRY(hipStreamSynchronize(st));
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

static void collect_profile(ifgf_plan* P) {
  const size_t per = events_per_apply(P);
  for (size_t a = 0; a + per <= P->ev_used; a += per) {
    hipEvent_t* ev = &P->ev_pool[a];
    HIP_TRY(hipEventSynchronize(ev[11]));
    for (int s = 0; s < 6; s++) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
      P->stage_ms[s] += ms;
    }
    P->prof_count++;
  }
  P->ev_used = 0;
}

// one device-resident application on stream st (used by ifgf_apply and GMRES)
int ifgf::apply_device(ifgf_plan* P, const zd* x, zd* y, hipStream_t st, bool fp64) {
  if (!dispatch(P, x, y, st, fp64)) return fail(P, IFGF_ERR_ARG, "no kernel instantiation for this (Ps, Pang)");
  // bound the number of outstanding event sets
  if (P->profile && P->ev_used >= (size_t)events_per_apply(P) * 4096) collect_profile(P);
  return IFGF_OK;
}

// nrhs device columns (stride ld) through the batched kernels in blocks of
// 8 / 4 / 2 columns where the plan has them, single applications otherwise.
// With two or more 8-column blocks and a second buffer set (nrhs_max >= 16)
// the 8-column blocks alternate between st and the second lane's stream, so
// one block's latency-bound level pushes overlap the other's bandwidth-bound
// near field, leaf operator and final sum; st waits for both lanes.
int ifgf::apply_columns(ifgf_plan* P, const zd* x, long long ld, zd* y, int nrhs, hipStream_t st, bool fp64) {
  const bool batched_ok = P->stored && P->rb_max >= 2;
  // lanes used: 1 + the extra lanes that are built and that the call's
  // 8-column blocks fill
  int nl = 1;
  if (batched_ok && P->rb_max == 8)
    while (nl < P->max_lanes && P->xlane[nl - 1].ready && nrhs >= 8 * (nl + 1)) nl++;
  for (int l = 1; l < nl; l++) {
    BatchLane& L = P->xlane[l - 1];
    HIP_TRY(hipEventRecord(L.ev_start, st));
    HIP_TRY(hipStreamWaitEvent(L.st, L.ev_start, 0));
  }
  int r = 0, blk = 0;
  while (r < nrhs) {
    int R = 1;
    if (batched_ok)
      for (int c : {8, 4, 2})
        if (c <= P->rb_max && c <= nrhs - r) {
          R = c;
          break;
        }
    const zd* xr = x + (size_t)r * ld;
    zd* yr = y + (size_t)r * ld;
    if (R > 1) {
      const int lane = R == 8 ? blk++ % nl : 0;
      if (!dispatch_block(P, R, xr, ld, yr, lane ? P->xlane[lane - 1].st : st, lane, fp64))
        return fail(P, IFGF_ERR_ARG, "no kernel instantiation");
    } else {
      const int rc = apply_device(P, xr, yr, st, fp64);
      if (rc) return rc;
    }
    r += R;
  }
  for (int l = 1; l < nl; l++) {
    BatchLane& L = P->xlane[l - 1];
    HIP_TRY(hipEventRecord(L.ev_done, L.st));
    HIP_TRY(hipStreamWaitEvent(st, L.ev_done, 0));
  }
  return IFGF_OK;
}

extern "C" {

int ifgf_apply(ifgf_plan* P, const double* x, double* y, int nrhs, long long ld, int flags, void* stream) {
  if (!P || !x || !y) return fail(P, IFGF_ERR_ARG, "null argument");
  if (!P->on_device) return fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long N = P->geo.N;
  if (nrhs < 1 || ld < N) return fail(P, IFGF_ERR_ARG, "nrhs < 1 or ld < n");
  try {
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    const bool fp64 = (flags & IFGF_APPLY_FP64) != 0;
    if (flags & IFGF_PTR_DEVICE) return ifgf::apply_columns(P, (const zd*)x, ld, (zd*)y, nrhs, st, fp64);
    const bool batched_ok = P->stored && P->rb_max >= 2;
    int r = 0;
    while (r < nrhs) {
      // column block: the widest batched width that fits, else one column
      int R = 1;
      if (batched_ok)
        for (int c : {8, 4, 2})
          if (c <= P->rb_max && c <= nrhs - r) {
            R = c;
            break;
          }
      const zd* xr = (const zd*)x + (size_t)r * ld;
      zd* yr = (zd*)y + (size_t)r * ld;
      // host columns: staged through xbuf / ybuf one block at a time
      for (int q = 0; q < R; q++)
        HIP_TRY(hipMemcpyAsync(P->xbuf.p + (size_t)q * N, xr + (size_t)q * ld, N * sizeof(zd), hipMemcpyHostToDevice,
                               st));
      if (R > 1) {
        if (!dispatch_block(P, R, P->xbuf.p, N, P->ybuf.p, st, 0, fp64))
          return fail(P, IFGF_ERR_ARG, "no kernel instantiation");
      } else {
        int rc = apply_device(P, P->xbuf.p, P->ybuf.p, st, fp64);
        if (rc) return rc;
      }
      for (int q = 0; q < R; q++)
        HIP_TRY(hipMemcpyAsync(yr + (size_t)q * ld, P->ybuf.p + (size_t)q * N, N * sizeof(zd), hipMemcpyDeviceToHost,
                               st));
      HIP_TRY(hipStreamSynchronize(st));
      r += R;
    }
  } catch (std::exception& e) {
    return fail(P, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_apply_graph(ifgf_plan* P, const double* x, double* y, int nrhs, long long ld, int flags, void* stream) {
  if (!P || !x || !y) return fail(P, IFGF_ERR_ARG, "null argument");
  if (!P->on_device) return fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  if (nrhs < 1 || ld < P->geo.N) return fail(P, IFGF_ERR_ARG, "nrhs < 1 or ld < n");
  if (P->profile) return fail(P, IFGF_ERR_ARG, "stage profiling is on (its events cannot be captured)");
  const int f = flags & IFGF_APPLY_FP64;
  try {
    if (!P->gexec || P->g_x != x || P->g_y != y || P->g_nrhs != nrhs || P->g_ld != ld || P->g_flags != f) {
      if (P->gexec) HIP_TRY(hipGraphExecDestroy(P->gexec));
      if (P->graph) HIP_TRY(hipGraphDestroy(P->graph));
      P->gexec = nullptr;
      P->graph = nullptr;
      // capture on the plan's stream: every application launch of nrhs
      // columns (kernels, and the fork / join events of the overlap path)
      HIP_TRY(hipStreamSynchronize(P->stream));
      HIP_TRY(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
      int rc = IFGF_OK;
      for (int r = 0; r < nrhs && rc == IFGF_OK; r++)
        rc = apply_device(P, (const zd*)x + (size_t)r * ld, (zd*)y + (size_t)r * ld, P->stream, f != 0);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(P->stream, &g);
      if (rc != IFGF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      HIP_TRY(ec);
      P->graph = g;
      HIP_TRY(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
      P->g_x = x;
      P->g_y = y;
      P->g_nrhs = nrhs;
      P->g_ld = ld;
      P->g_flags = f;
    }
    HIP_TRY(hipGraphLaunch(P->gexec, stream ? (hipStream_t)stream : P->stream));
  } catch (std::exception& e) {
    return fail(P, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

}  // extern "C"

// A group of plans (obstacles) applied as one captured graph.
// ---- obstacle groups: every stage of every plan in one launch --------------
// A fused launch runs the blocks of all plans of a class (same (Ps, Pang) and
// precision; the stored-path kernels do not depend on the formulation) in one
// grid: block b finds its plan i from the block offsets boff (the largest i
// with boff[i] <= b), loads that plan's arguments and runs the single-plan
// kernel's body with block index b - boff[i]. A group application is then
// gather, near field, leaf operator, one push launch per tree depth (depth d
// of a plan with D levels is its level D - d) and the finalize (one launch per
// finalize lane count in the class), all captured into one HIP graph.
namespace {

struct GatherArgs {
  int N;
  const zd* x;
  const double* w;
  const int* perm;
  zd* rho;
};
struct NearArgs {
  int N, nitems;
  const NearItem* items;
  const int4* nf_seg;
  const zd* nf_val;
  const zd* rho;
  zd* npart;
};
template <typename Z>
struct LeafArgs {
  int nitems;
  const LeafItem* items;
  const Z* lf;
  const zd* rho;
  Z* lpart;
};
template <typename Z>
struct FinArgs {
  int N, Np;
  double ident;
  const zd* x;
  const long long* corr_ptr;
  const int* corr_patch;
  const zd* corr_w;
  const int* inv;
  const zd* sol;
  const int* nio;
  const zd* npart;
  const long long* it_ptr;
  const int* it_slot;
  const Z* slots;
  zd* y;
};

__device__ inline int group_plan(const int* __restrict__ boff, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (boff[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_group_gather(const GatherArgs* __restrict__ a, const int* __restrict__ boff,
                                                      int n) {
  const int b = xcd_block(), i = group_plan(boff, n, b);
  const GatherArgs A = a[i];
  gather_body(b - boff[i], A.N, A.x, A.w, A.perm, A.rho);
}
__global__ void __launch_bounds__(256) k_group_near(const NearArgs* __restrict__ a, const int* __restrict__ boff,
                                                    int n) {
  const int b = xcd_block(), i = group_plan(boff, n, b);
  const NearArgs A = a[i];
  near_items_body<false>(b - boff[i], A.N, A.nitems, A.items, A.nf_seg, A.nf_val, A.rho, A.npart);
}
template <int P, typename Z>
__global__ void __launch_bounds__(256) k_group_leaf(const LeafArgs<Z>* __restrict__ a, const int* __restrict__ boff,
                                                    int n) {
  const int b = xcd_block(), i = group_plan(boff, n, b);
  const LeafArgs<Z> A = a[i];
  leaf_items_body<P, Z>(b - boff[i], A.nitems, A.items, A.lf, A.rho, A.lpart);
}
template <int Ps, int Pang, int EPT, typename R>
__global__ void __launch_bounds__(256) k_group_push(const PushArgsT<R>* __restrict__ a, const int* __restrict__ boff,
                                                    int n) {
  __shared__ PushShared<Ps, Pang, R> sh;
  const int b = xcd_block(), i = group_plan(boff, n, b);
  const PushArgsT<R> A = a[i];
  level_push_body<Ps, Pang, EPT, R>(b - boff[i], A, sh);
}
template <int S, typename Z>
__global__ void __launch_bounds__(256) k_group_fin(const FinArgs<Z>* __restrict__ a, const int* __restrict__ boff,
                                                   int n) {
  const int b = xcd_block(), i = group_plan(boff, n, b);
  const FinArgs<Z> A = a[i];
  finalize_body<S, Z>(b - boff[i], A.N, A.Np, A.ident, A.x, A.corr_ptr, A.corr_patch, A.corr_w, A.inv, A.sol, A.nio,
                      A.npart, A.it_ptr, A.it_slot, A.slots, A.y, 0, nullptr);
}

// One fused stage: per-plan arguments and block offsets on the device.
struct Stage {
  DevBuf<unsigned char> args;
  DevBuf<int> boff;
  int n = 0, blocks = 0;
  template <typename A>
  void set(const std::vector<A>& a, const std::vector<int>& nb) {
    n = (int)a.size();
    std::vector<int> off(n + 1, 0);
    for (int i = 0; i < n; i++) off[i + 1] = off[i] + nb[i];
    blocks = off[n];
    std::vector<unsigned char> raw(sizeof(A) * a.size());
    if (!a.empty()) std::memcpy(raw.data(), a.data(), raw.size());
    args.upload(raw);
    boff.upload(off);
  }
  template <typename A>
  const A* p() const {
    return reinterpret_cast<const A*>(args.p);
  }
};

// Plans of one class and its fused stages.
struct GroupClass {
  int Ps = 0, Pang = 0, ept = 1;
  bool f32 = false;
  std::vector<int> plans;  // indices into the group's plans
  Stage gather, near, leaf;
  std::vector<Stage> push;  // per depth
  Stage fin[3];             // finalize lanes 4, 8, 16
};

}  // namespace

struct ifgf_group {
  std::vector<ifgf_plan*> plans;
  std::vector<GroupClass> classes;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;
  ~ifgf_group() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <typename R>
PushArgsT<R> push_args_of(ifgf_plan* P, int li) {
  const int D = P->tree.D;
  PushArgsT<R> a{};
  if constexpr (std::is_same<R, float>::value) {
    a = PushArgsT<R>{nullptr, nullptr, P->push_geo32.p, nullptr, P->push_coff.p, P->push_cidx.p,
                     P->leaf_item_off.p, P->lpart32.p, P->tree.lv[D].cone_off, P->ifac_off.p, P->ifac_tab32.p,
                     P->tabs32.p, P->pbuf32.p, 0};
  } else {
    const bool direct = P->push_geo_b.n > 0;
    a = PushArgsT<R>{nullptr, direct ? nullptr : P->push_einfo.p, direct ? P->push_geo_b.p : P->push_igeo.p,
                     direct ? nullptr : P->push_ptab.p, P->push_coff.p, P->push_cidx.p,
                     P->leaf_item_off.p, P->lpart.p, P->tree.lv[D].cone_off, P->ifac_off.p, P->ifac_tab.p,
                     P->tabs.p, P->pbuf.p, 0};
  }
  a.tasks = P->push_tasks.p + P->push_level_off[li];
  a.leaf = li == 0;
  return a;
}

// Builds the fused stages of class C (plans of precision R, P = Ps * Pang).
template <int P_, typename R>
void build_class(GroupClass& C, const std::vector<ifgf_plan*>& plans, const double* const* x, double* const* y) {
  using Z = typename complex_of<R>::type;
  constexpr bool f32 = std::is_same<R, float>::value;
  std::vector<GatherArgs> ga;
  std::vector<NearArgs> na;
  std::vector<LeafArgs<Z>> la;
  std::vector<int> gb, nbk, lb;
  int maxd = 0;
  C.ept = 1;
  for (int q : C.plans) {
    ifgf_plan* P = plans[q];
    const int N = (int)P->tree.N;
    ga.push_back(GatherArgs{N, (const zd*)x[q], P->w.p, P->perm.p, P->rho.p});
    gb.push_back((N + 255) / 256);
    const int M = (int)P->near_items.n;
    na.push_back(NearArgs{(int)P->tree.M, M, P->near_items.p, P->nf_seg.p, P->nf_val.p, P->rho.p, P->npart.p});
    nbk.push_back((M + 3) / 4);
    const long long L = P->leaf_items.n;
    la.push_back(LeafArgs<Z>{(int)L, P->leaf_items.p, f32 ? (const Z*)P->lf_val32.p : (const Z*)P->lf_val.p,
                             P->rho.p, f32 ? (Z*)P->lpart32.p : (Z*)P->lpart.p});
    lb.push_back((int)((L * P_ * kLeafChunk + 255) / 256));
    maxd = std::max(maxd, P->tree.D - 1);
    C.ept = std::max(C.ept, P->push_emax / 256);
  }
  C.gather.set(ga, gb);
  C.near.set(na, nbk);
  C.leaf.set(la, lb);
  C.push.resize(maxd);
  for (int li = 0; li < maxd; li++) {
    std::vector<PushArgsT<R>> pa;
    std::vector<int> pb;
    for (int q : C.plans) {
      ifgf_plan* P = plans[q];
      const bool has = li + 1 < P->tree.D;
      pa.push_back(has ? push_args_of<R>(P, li) : PushArgsT<R>{});
      pb.push_back(has ? P->push_level_off[li + 1] - P->push_level_off[li] : 0);
    }
    C.push[li].set(pa, pb);
  }
  for (int si = 0; si < 3; si++) {
    const int S = 4 << si;
    std::vector<FinArgs<Z>> fa;
    std::vector<int> fb;
    for (int q : C.plans) {
      ifgf_plan* P = plans[q];
      const bool mine = P->final_lanes == S;
      const int NT = (int)P->tree.M;
      const double ident = P->prm.formulation == kSingleLayer ? 0.0 : 0.5;
      fa.push_back(FinArgs<Z>{NT, P->geo.Np, ident, (const zd*)x[q], P->corr_ptr.p, P->corr_patch.p, P->corr_w.p,
                              P->inv.p, P->sol.p, P->near_item_off.p, P->npart.p, P->it_ptr.p, P->it_slot.p,
                              f32 ? (const Z*)P->pbuf32.p : (const Z*)P->pbuf.p, (zd*)y[q]});
      fb.push_back(mine ? (NT + 256 / S - 1) / (256 / S) : 0);
    }
    C.fin[si].set(fa, fb);
  }
}

template <int Ps, int Pang, int EPT, typename R>
void launch_class_push(const GroupClass& C, hipStream_t st) {
  for (const Stage& S : C.push)
    if (S.blocks)
      hipLaunchKernelGGL((k_group_push<Ps, Pang, EPT, R>), dim3(S.blocks), dim3(256), 0, st, S.p<PushArgsT<R>>(),
                         S.boff.p, S.n);
}

template <int Ps, int Pang, typename R>
void launch_class(const GroupClass& C, hipStream_t st) {
  using Z = typename complex_of<R>::type;
  constexpr int P = Ps * Pang;
  if (C.gather.blocks)
    hipLaunchKernelGGL(k_group_gather, dim3(C.gather.blocks), dim3(256), 0, st, C.gather.p<GatherArgs>(),
                       C.gather.boff.p, C.gather.n);
  if (C.near.blocks)
    hipLaunchKernelGGL(k_group_near, dim3(C.near.blocks), dim3(256), 0, st, C.near.p<NearArgs>(), C.near.boff.p,
                       C.near.n);
  if (C.leaf.blocks)
    hipLaunchKernelGGL((k_group_leaf<P, Z>), dim3(C.leaf.blocks), dim3(256), 0, st, C.leaf.p<LeafArgs<Z>>(),
                       C.leaf.boff.p, C.leaf.n);
  switch (C.ept) {
    case 1: launch_class_push<Ps, Pang, 1, R>(C, st); break;
    case 2: launch_class_push<Ps, Pang, 2, R>(C, st); break;
    case 4: launch_class_push<Ps, Pang, 4, R>(C, st); break;
    default: launch_class_push<Ps, Pang, 8, R>(C, st); break;
  }
  if (C.fin[0].blocks)
    hipLaunchKernelGGL((k_group_fin<4, Z>), dim3(C.fin[0].blocks), dim3(256), 0, st, C.fin[0].p<FinArgs<Z>>(),
                       C.fin[0].boff.p, C.fin[0].n);
  if (C.fin[1].blocks)
    hipLaunchKernelGGL((k_group_fin<8, Z>), dim3(C.fin[1].blocks), dim3(256), 0, st, C.fin[1].p<FinArgs<Z>>(),
                       C.fin[1].boff.p, C.fin[1].n);
  if (C.fin[2].blocks)
    hipLaunchKernelGGL((k_group_fin<16, Z>), dim3(C.fin[2].blocks), dim3(256), 0, st, C.fin[2].p<FinArgs<Z>>(),
                       C.fin[2].boff.p, C.fin[2].n);
}

// (Ps, Pang) dispatch of the class builders and launchers
bool class_dispatch(GroupClass& C, const std::vector<ifgf_plan*>& plans, const double* const* x, double* const* y,
                    hipStream_t st, bool build) {
#define IFGF_GCASE(A, B)                                                          \
  if (C.Ps == A && C.Pang == B) {                                                 \
    if (build) {                                                                  \
      if (C.f32) build_class<A * B, float>(C, plans, x, y);                       \
      else build_class<A * B, double>(C, plans, x, y);                            \
    } else {                                                                      \
      if (C.f32) launch_class<A, B, float>(C, st);                                \
      else launch_class<A, B, double>(C, st);                                     \
    }                                                                             \
    return true;                                                                  \
  }
  IFGF_GCASE(5, 5)
  IFGF_GCASE(3, 5)
  IFGF_GCASE(5, 7)
  IFGF_GCASE(7, 7)
  IFGF_GCASE(9, 9)
  IFGF_GCASE(11, 11)
#undef IFGF_GCASE
  return false;
}

}  // namespace

extern "C" {

int ifgf_group_create(ifgf_plan* const* plans, int nplans, const double* const* x, double* const* y,
                      ifgf_group** out) {
  if (!plans || nplans < 1 || !x || !y || !out) return fail(nullptr, IFGF_ERR_ARG, "invalid argument");
  *out = nullptr;
  for (int i = 0; i < nplans; i++) {
    if (!plans[i] || !x[i] || !y[i]) return fail(nullptr, IFGF_ERR_ARG, "null plan or vector");
    if (!plans[i]->on_device) return fail(nullptr, IFGF_ERR_NO_DEVICE, "plan was built host-only");
    if (!plans[i]->stored || plans[i]->field)
      return fail(nullptr, IFGF_ERR_ARG, "group plans need stored operators (not on the fly, not a field)");
    for (int j = 0; j < i; j++)
      if (plans[j] == plans[i]) return fail(nullptr, IFGF_ERR_ARG, "a plan appears twice in the group");
  }
  std::unique_ptr<ifgf_group> G(new ifgf_group);
  try {
    G->plans.assign(plans, plans + nplans);
    for (int i = 0; i < nplans; i++) HIP_TRY(hipStreamSynchronize(plans[i]->stream));
    // classes: plans of equal (Ps, Pang) and precision, in order of first appearance
    for (int i = 0; i < nplans; i++) {
      const ifgf_plan* P = plans[i];
      GroupClass* C = nullptr;
      for (auto& c : G->classes)
        if (c.Ps == P->prm.Ps && c.Pang == P->prm.Pang && c.f32 == P->mixed) C = &c;
      if (!C) {
        G->classes.emplace_back();
        C = &G->classes.back();
        C->Ps = P->prm.Ps;
        C->Pang = P->prm.Pang;
        C->f32 = P->mixed;
      }
      C->plans.push_back(i);
    }
    for (auto& C : G->classes)
      if (!class_dispatch(C, G->plans, x, y, nullptr, true))
        return fail(nullptr, IFGF_ERR_ARG, "no kernel instantiation for a plan's (Ps, Pang)");
    HIP_TRY(hipStreamCreate(&G->stream));
    HIP_TRY(hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal));
    for (auto& C : G->classes) class_dispatch(C, G->plans, x, y, G->stream, false);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(G->stream, &g);
    HIP_TRY(ec);
    HIP_TRY(ee);
    G->graph = g;
    HIP_TRY(hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0));
  } catch (std::exception& e) {
    return fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  *out = G.release();
  return IFGF_OK;
}

int ifgf_group_launch(ifgf_group* G, void* stream) {
  if (!G) return fail(nullptr, IFGF_ERR_ARG, "null group");
  try {
    HIP_TRY(hipGraphLaunch(G->exec, stream ? (hipStream_t)stream : G->stream));
  } catch (std::exception& e) {
    return fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_group_synchronize(ifgf_group* G) {
  if (!G) return fail(nullptr, IFGF_ERR_ARG, "null group");
  try {
    HIP_TRY(hipStreamSynchronize(G->stream));
  } catch (std::exception& e) {
    return fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_autotune_cone_order(const ifgf_surface* surface, const ifgf_params* params, double tol, int* Ps, int* Pang,
                             double* err) {
  if (!surface || !params || !Ps || !Pang || !err || !(tol > 0.0)) return fail(nullptr, IFGF_ERR_ARG, "invalid argument");
  static const int cand[][2] = {{3, 5}, {5, 5}, {5, 7}, {7, 7}, {9, 9}, {11, 11}};
  double best = 1e300;
  int bp = 0;
  for (int c = 0; c < 6; c++) {
    ifgf_params prm = *params;
    prm.Ps = cand[c][0];
    prm.Pang = cand[c][1];
    prm.nrhs_max = 1;
    ifgf_plan* P = nullptr;
    int rc = ifgf_plan_create(surface, &prm, &P);
    if (rc != IFGF_OK) return rc;
    const long long N = P->geo.N;
    // seeded complex density (splitmix64 uniforms in [-1, 1)) and a spread of
    // at most 256 targets: the interpolation error against the direct sum
    std::vector<double> x(2 * N), y(2 * N);
    for (long long i = 0; i < 2 * N; i++) {
      uint64_t z = 0x5eedull + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    const long long nt = std::min<long long>(N, 256);
    std::vector<long long> tg(nt);
    for (long long t = 0; t < nt; t++) tg[t] = t * N / nt;
    std::vector<double> yd(2 * nt);
    rc = ifgf_apply(P, x.data(), y.data(), 1, N, 0, nullptr);
    if (rc == IFGF_OK) rc = ifgf_apply_direct(P, x.data(), tg.data(), nt, yd.data());
    ifgf_plan_destroy(P);
    if (rc != IFGF_OK) return rc;
    double num = 0.0, den = 0.0;
    for (long long t = 0; t < nt; t++) {
      const double dr = y[2 * tg[t]] - yd[2 * t], di = y[2 * tg[t] + 1] - yd[2 * t + 1];
      num += dr * dr + di * di;
      den += yd[2 * t] * yd[2 * t] + yd[2 * t + 1] * yd[2 * t + 1];
    }
    const double e = std::sqrt(num / den);
    if (e < best) {
      best = e;
      bp = c;
    }
    if (e <= tol) break;
  }
  *Ps = cand[bp][0];
  *Pang = cand[bp][1];
  *err = best;
  return best <= tol ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
}

int ifgf_group_stats(const ifgf_group* G, long long out[4]) {
  if (!G || !out) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  long long launches = 0, blocks = 0;
  for (const auto& C : G->classes) {
    const Stage* st[] = {&C.gather, &C.near, &C.leaf, &C.fin[0], &C.fin[1], &C.fin[2]};
    for (const Stage* s : st) {
      launches += s->blocks > 0;
      blocks += s->blocks;
    }
    for (const auto& s : C.push) {
      launches += s.blocks > 0;
      blocks += s.blocks;
    }
  }
  out[0] = (long long)G->plans.size();
  out[1] = (long long)G->classes.size();
  out[2] = launches;
  out[3] = blocks;
  return IFGF_OK;
}

void ifgf_group_destroy(ifgf_group* G) { delete G; }

int ifgf_apply_direct(ifgf_plan* P, const double* x, const long long* targets, long long nt, double* out) {
  if (!P || !x || !targets || !out) return fail(P, IFGF_ERR_ARG, "null argument");
  if (!P->on_device) return fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long N = P->geo.N;
  const int Np = P->geo.Np;
  try {
    std::vector<int> tg(nt), ex((size_t)nt * 4);
    for (long long t = 0; t < nt; t++) {
      if (targets[t] < 0 || targets[t] >= N) return fail(P, IFGF_ERR_ARG, "target index out of range");
      tg[t] = (int)targets[t];
      const int s = P->tree.inv[tg[t]];
      for (int q = 0; q < 4; q++) ex[t * 4 + q] = P->excl[(size_t)s * 4 + q];
    }
    std::vector<zd> rho(N);
    const zd* xv = (const zd*)x;
    for (long long i = 0; i < N; i++) rho[i] = cscale(xv[i], P->geo.w[i]);
    DevBuf<int> dt, de;
    DevBuf<double> ox, oy, onx, ony;
    DevBuf<zd> dr, dout;
    dt.upload(tg);
    de.upload(ex);
    ox.upload(P->geo.x);
    oy.upload(P->geo.y);
    onx.upload(P->geo.nx);
    ony.upload(P->geo.ny);
    dr.upload(rho);
    dout.alloc(nt);
    const int nb = (int)((nt + 255) / 256);
    const double k = P->prm.k, eta = P->prm.eta;
    switch (P->prm.formulation) {
      case kSingleLayer:
        hipLaunchKernelGGL(k_direct<kSingleLayer>, dim3(nb), dim3(256), 0, P->stream, (int)nt, dt.p, de.p,
                           (int)N, Np, ox.p, oy.p, onx.p, ony.p, dr.p, dout.p, k, eta);
        break;
      case kDoubleLayer:
        hipLaunchKernelGGL(k_direct<kDoubleLayer>, dim3(nb), dim3(256), 0, P->stream, (int)nt, dt.p, de.p,
                           (int)N, Np, ox.p, oy.p, onx.p, ony.p, dr.p, dout.p, k, eta);
        break;
      default:
        hipLaunchKernelGGL(k_direct<kCombined>, dim3(nb), dim3(256), 0, P->stream, (int)nt, dt.p, de.p,
                           (int)N, Np, ox.p, oy.p, onx.p, ony.p, dr.p, dout.p, k, eta);
    }
    HIP_TRY(hipGetLastError());
    std::vector<zd> res(nt);
    HIP_TRY(hipMemcpyAsync(res.data(), dout.p, nt * sizeof(zd), hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    const double ident = P->prm.formulation == kSingleLayer ? 0.0 : 0.5;
    zd* o = (zd*)out;
    for (long long t = 0; t < nt; t++) {
      const long long i = tg[t];
      zd acc = cscale(xv[i], ident);
      for (long long e = P->corr.ptr[i]; e < P->corr.ptr[i + 1]; e++)
        for (int q = 0; q < Np; q++)
          cfma(acc, P->corr.w[(size_t)e * Np + q], xv[(size_t)P->corr.patch[e] * Np + q]);
      o[t] = cadd(acc, res[t]);
    }
  } catch (std::exception& e) {
    return fail(P, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_propagate(ifgf_plan* P, const double* tx, const double* ty, long long nt, const double* dens,
                   double* out) {
  if (!P || !tx || !ty || !dens || !out) return fail(P, IFGF_ERR_ARG, "null argument");
  if (!P->on_device) return fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only");
  const long long N = P->geo.N;
  try {
    std::vector<zd> rho(N);
    const zd* dv = (const zd*)dens;
    for (long long i = 0; i < N; i++) rho[i] = cscale(dv[i], P->geo.w[i]);
    DevBuf<double> dx, dy, ox, oy, onx, ony;
    DevBuf<zd> dr, dout;
    dx.upload(std::vector<double>(tx, tx + nt));
    dy.upload(std::vector<double>(ty, ty + nt));
    ox.upload(P->geo.x);
    oy.upload(P->geo.y);
    onx.upload(P->geo.nx);
    ony.upload(P->geo.ny);
    dr.upload(rho);
    dout.alloc(nt);
    const int nb = (int)((nt + 255) / 256);
    const double k = P->prm.k, eta = P->prm.eta;
    switch (P->prm.formulation) {
      case kSingleLayer:
        hipLaunchKernelGGL(k_propagate<kSingleLayer>, dim3(nb), dim3(256), 0, P->stream, (int)nt, dx.p, dy.p,
                           (int)N, ox.p, oy.p, onx.p, ony.p, dr.p, dout.p, k, eta);
        break;
      case kDoubleLayer:
        hipLaunchKernelGGL(k_propagate<kDoubleLayer>, dim3(nb), dim3(256), 0, P->stream, (int)nt, dx.p, dy.p,
                           (int)N, ox.p, oy.p, onx.p, ony.p, dr.p, dout.p, k, eta);
        break;
      default:
        hipLaunchKernelGGL(k_propagate<kCombined>, dim3(nb), dim3(256), 0, P->stream, (int)nt, dx.p, dy.p,
                           (int)N, ox.p, oy.p, onx.p, ony.p, dr.p, dout.p, k, eta);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout.p, nt * sizeof(zd), hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
  } catch (std::exception& e) {
    return fail(P, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

}  // extern "C"

struct ifgf_field {
  std::unique_ptr<ifgf_plan> P;
};

extern "C" {

int ifgf_field_create(const ifgf_plan* src, const double* tx, const double* ty, long long nt, int nlevels, int Ps,
                      int Pang, int precision, ifgf_field** out) {
  if (!src || !tx || !ty || !out) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  *out = nullptr;
  if (nt < 1 || nt > (1ll << 30)) return fail(nullptr, IFGF_ERR_ARG, "target count out of range");
  if (nlevels < 3) return fail(nullptr, IFGF_ERR_LEVELS, "IFGF accelerator requires at least 3 levels.");
  if (Ps < 2 || Ps > 16 || Pang < 2 || Pang > 16) return fail(nullptr, IFGF_ERR_ARG, "Ps and Pang must lie in [2, 16]");
  if (precision != IFGF_FP64 && precision != IFGF_MIXED) return fail(nullptr, IFGF_ERR_ARG, "unsupported precision mode");
  std::unique_ptr<ifgf_field> F(new ifgf_field);
  F->P.reset(new ifgf_plan);
  ifgf_plan* P = F->P.get();
  P->field = true;
  P->prm = src->prm;
  P->prm.nlevels = nlevels;
  P->prm.Ps = Ps;
  P->prm.Pang = Pang;
  P->prm.precision = precision;
  P->prm.nrhs_max = 1;
  // a tuned plan's field is tuned too; a plan with the device tree builds the field's tree there
  P->prm.flags = src->prm.flags & (IFGF_PLAN_AUTOTUNE | IFGF_PLAN_DEVICE_TREE);
  P->geo = src->geo;
  std::string err;
  try {
    const Geometry& g = P->geo;
    const bool dev_tree = (P->prm.flags & IFGF_PLAN_DEVICE_TREE) && src->on_device;
    if (!build_tree(g.x.data(), g.y.data(), g.nx.data(), g.ny.data(), g.N, nlevels, P->prm.k, Ps, Pang, P->tree, err,
                    nullptr, tx, ty, nt, dev_tree ? relevant_cones_device : nullptr,
                    dev_tree ? tree_groups_device : nullptr))
      return fail(nullptr, IFGF_ERR_GEOMETRY, err);
    P->n_near = P->tree.n_near_t;
    build_cone_tables(P);
    if (src->on_device) {  // a field of a host-only plan holds the tables only
      build_device_side(P);
      if (!build_stored_any(P)) return fail(nullptr, IFGF_ERR_ARG, "no kernel instantiation for this (Ps, Pang)");
      if (P->prm.flags & IFGF_PLAN_AUTOTUNE) autotune(P);
    }
  } catch (std::exception& e) {
    return fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  *out = F.release();
  return IFGF_OK;
}

int ifgf_field_apply(ifgf_field* F, const double* density, double* out, int flags, void* stream) {
  if (!F || !density || !out) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  ifgf_plan* P = F->P.get();
  if (!P->on_device) return fail(P, IFGF_ERR_NO_DEVICE, "field of a host-only plan");
  const long long N = P->geo.N, M = P->tree.M;
  try {
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    const bool fp64 = (flags & IFGF_APPLY_FP64) != 0;
    if (flags & IFGF_PTR_DEVICE) return apply_device(P, (const zd*)density, (zd*)out, st, fp64);
    HIP_TRY(hipMemcpyAsync(P->xbuf.p, density, N * sizeof(zd), hipMemcpyHostToDevice, st));
    const int rc = apply_device(P, P->xbuf.p, P->ybuf.p, st, fp64);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, P->ybuf.p, M * sizeof(zd), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  } catch (std::exception& e) {
    return fail(P, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

int ifgf_field_stats(const ifgf_field* F, long long out[8]) {
  if (!F || !out) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  const ifgf_plan* P = F->P.get();
  out[0] = P->tree.M;
  out[1] = P->tree.n_near_t;
  out[2] = P->tree.n_leaf;
  out[3] = P->tree.n_prop;
  out[4] = P->tree.n_interp;
  out[5] = P->tree.n_cones;
  out[6] = P->tree.nlev;
  out[7] = P->stored_bytes;
  return IFGF_OK;
}

void ifgf_field_destroy(ifgf_field* F) { delete F; }

int ifgf_hankel01(const double* x, long long n, double* h0, double* h1) {
  if (!x || !h0 || !h1 || n < 0) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  try {
    DevBuf<double> dx;
    DevBuf<zd> d0, d1;
    dx.upload(std::vector<double>(x, x + n));
    d0.alloc(n);
    d1.alloc(n);
    hipLaunchKernelGGL(k_hankel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, dx.p, d0.p, d1.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(h0, d0.p, n * sizeof(zd), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(h1, d1.p, n * sizeof(zd), hipMemcpyDeviceToHost));
  } catch (std::exception& e) {
    return fail(nullptr, IFGF_ERR_HIP, e.what());
  }
  return IFGF_OK;
}

}  // extern "C"

// ifgf_propagate_noalloc's scratch: the intensity-weighted density of the
// plan's points in Morton order, 2 doubles per point, for at most
// kNoAllocMaxPoints points (16 MiB on every device the library runs on, held
// until the process ends). Host code never passes this array's name to a
// kernel: its device address comes from hipGetSymbolAddress, once per device.
namespace {
constexpr long long kNoAllocMaxPoints = 1ll << 20;
__device__ __align__(256) double scratch[2 * kNoAllocMaxPoints];
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of scratch (looked up on the device's first call).
int scratch_address(ifgf::zd** out, std::string& err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return IFGF_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return IFGF_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = (ifgf::zd*)g_scratch_ptr[dev];
  return IFGF_OK;
}
}  // namespace

extern "C" {

int ifgf_propagate_noalloc(ifgf_plan* P, const double* tx, const double* ty, long long nt, const double* density,
                           double* out, void* stream) {
  if (!P || (nt > 0 && (!tx || !ty || !out)) || !density) return fail(P, IFGF_ERR_ARG, "null argument");
  if (nt < 0 || nt > 2147483647ll) return fail(P, IFGF_ERR_ARG, "target count out of range");
  if (!P->on_device || P->field) return fail(P, IFGF_ERR_NO_DEVICE, "plan was built host-only or is a field");
  const long long N = P->geo.N;
  if (N > kNoAllocMaxPoints)
    return fail(P, IFGF_ERR_ARG, "plan has " + std::to_string(N) + " points; ifgf_propagate_noalloc's scratch holds " +
                                     std::to_string(kNoAllocMaxPoints));
  if (nt == 0) return IFGF_OK;
  ifgf::zd* rho = nullptr;
  std::string err;
  if (const int rc = scratch_address(&rho, err)) return fail(P, rc, err);
  hipStream_t st = stream ? (hipStream_t)stream : P->stream;
  // every scratch entry the direct sum reads is written by the gather first
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, (int)N, (const zd*)density, P->w.p,
                     P->perm.p, rho);
  const int nb = (int)((nt + 255) / 256);
  const double k = P->prm.k, eta = P->prm.eta;
  const zd* r = rho;
  zd* o = (zd*)out;
  switch (P->prm.formulation) {
    case kSingleLayer:
      hipLaunchKernelGGL(k_propagate<kSingleLayer>, dim3(nb), dim3(256), 0, st, (int)nt, tx, ty, (int)N, P->px.p,
                         P->py.p, P->pnx.p, P->pny.p, r, o, k, eta);
      break;
    case kDoubleLayer:
      hipLaunchKernelGGL(k_propagate<kDoubleLayer>, dim3(nb), dim3(256), 0, st, (int)nt, tx, ty, (int)N, P->px.p,
                         P->py.p, P->pnx.p, P->pny.p, r, o, k, eta);
      break;
    default:
      hipLaunchKernelGGL(k_propagate<kCombined>, dim3(nb), dim3(256), 0, st, (int)nt, tx, ty, (int)N, P->px.p,
                         P->py.p, P->pnx.p, P->pny.p, r, o, k, eta);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(P, IFGF_ERR_HIP, hipGetErrorString(e));
  return IFGF_OK;
}

int ifgf_curve_eval(int curve_type, const double* params, int nparams, const double* t, long long n, double* x,
                    double* y, double* xp, double* yp) {
  if (!params || (n > 0 && !t) || n < 0) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  Curve c;
  std::string err;
  if (!make_curve(curve_type, params, nparams, c, err)) return fail(nullptr, IFGF_ERR_ARG, err);
  for (long long i = 0; i < n; i++) {
    if (x) x[i] = c.x(t[i]);
    if (y) y[i] = c.y(t[i]);
    if (xp) xp[i] = c.xp(t[i]);
    if (yp) yp[i] = c.yp(t[i]);
  }
  return IFGF_OK;
}

int ifgf_curve_diff(int curve_type, const double* params, int nparams, const double* a, const double* d,
                    long long n, double* dx, double* dy) {
  if (!params || n < 0 || (n > 0 && (!a || !d || !dx || !dy))) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  Curve c;
  std::string err;
  if (!make_curve(curve_type, params, nparams, c, err)) return fail(nullptr, IFGF_ERR_ARG, err);
  for (long long i = 0; i < n; i++) c.diff(a[i], d[i], dx[i], dy[i]);
  return IFGF_OK;
}

int ifgf_plan_profile(ifgf_plan* P, int enable) {
  if (!P) return fail(nullptr, IFGF_ERR_ARG, "null plan");
  P->profile = enable != 0;
  P->ev_used = 0;
  for (double& v : P->stage_ms) v = 0.0;
  P->prof_count = 0;
  return IFGF_OK;
}

int ifgf_plan_setup_times(const ifgf_plan* P, double seconds[8]) {
  if (!P || !seconds) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  for (int q = 0; q < 8; q++) seconds[q] = P->setup_s[q];
  return IFGF_OK;
}

int ifgf_plan_stage_times(ifgf_plan* P, double times_ms[6], long long* count) {
  if (!P || !times_ms) return fail(nullptr, IFGF_ERR_ARG, "null argument");
  try {
    collect_profile(P);
  } catch (std::exception& e) {
    return fail(P, IFGF_ERR_HIP, e.what());
  }
  for (int s = 0; s < 6; s++) times_ms[s] = P->stage_ms[s];
  if (count) *count = P->prof_count;
  return IFGF_OK;
}

int ifgf_gmres_host(ifgf_apply_fn apply, void* ctx, long long n, const double* rhs, double* x, int restart,
                    double tol, int maxit, int* iters, double* err) {
  if (!apply || !rhs || !x || n < 1 || restart < 1 || !iters || !err)
    return fail(nullptr, IFGF_ERR_ARG, "invalid argument");
  const bool ok = householder_gmres(
      [&](const std::complex<double>* a, std::complex<double>* b) { apply(ctx, (const double*)a, (double*)b); },
      n, (const std::complex<double>*)rhs, (std::complex<double>*)x, restart, tol, maxit, *iters, *err);
  return ok ? IFGF_OK : IFGF_ERR_NOT_CONVERGED;
}

}  // extern "C"
