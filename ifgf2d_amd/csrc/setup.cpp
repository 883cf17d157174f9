// Host setup of the IFGF forward map (geometry, RP correction, tree tables).
// See setup.h for the data layout. Reference citations are file:line in
// manuelarturosantana/IFGF2D.
#include "setup.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <unordered_map>

#include "cone_grid.h"
#include "green.h"

namespace ifgf {

// ---------------------------------------------------------------- curves ---
bool make_curve(int type, const double* prm, int nprm, Curve& c, std::string& err) {
  c = Curve();
  c.type = type;
  if (type == 0 && nprm >= 3) {
    c.cx = prm[0];
    c.cy = prm[1];
    c.R = prm[2];
    return true;
  }
  if (type == 1 && nprm >= 4) {
    c.A = prm[0];
    c.B = prm[1];
    c.C = prm[2];
    c.D = prm[3];
    return true;
  }
  if (type == 2 && nprm >= 4) {
    c.cx = prm[0];
    c.cy = prm[1];
    c.R = prm[2];
    const int nm = (int)prm[3];
    if (nm < 0 || nprm < 4 + 2 * nm) {
      err = "star curve: parameter list shorter than 4 + 2*nmodes";
      return false;
    }
    if (nm > kMaxStarModes) {
      err = "star curve: more than 64 modes";
      return false;
    }
    c.nm = nm;
    for (int m = 0; m < nm; m++) {
      c.am[m] = prm[4 + m];
      c.bm[m] = prm[4 + nm + m];
    }
    return true;
  }
  err = "unknown curve type or too few parameters";
  return false;
}

// ------------------------------------------------------------ quadrature ---
// Fejer's first rule, utils/Quadrature.hpp:25-43.
static void fejer(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; i++) {
    x[i] = std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * n));
    double s = 0.0;
    for (int j = 1; j <= n / 2; j++) s += std::cos(j * M_PI * (2.0 * i + 1.0) / n) / (4.0 * j * j - 1.0);
    w[i] = (2.0 / n) * (1.0 - 2.0 * s);
  }
}

// -------------------------------------------------------------- geometry ---
bool build_geometry(const Curve& c, const double* lims, int npatches, int Np, double delta,
                    Geometry& g, std::string& err, NsSearch ns_search) {
  if (npatches < 3 || Np < 2 || Np > 64) {
    err = "need at least 3 patches and 2 <= Np <= 64";
    return false;
  }
  g = Geometry();
  g.Np = Np;
  g.npatches = npatches;
  g.N = (long long)npatches * Np;
  fejer(Np, g.fejer_nodes, g.fejer_weights);
  g.t1.resize(npatches);
  g.t2.resize(npatches);
  for (auto* v : {&g.x, &g.y, &g.nx, &g.ny, &g.w, &g.t}) v->resize(g.N);
  // ForwardMap.tpp:75-157: node t = ab2cd(fejer_j, -1, 1, t1, t2), normal (y', -x'),
  // intensity factor (t2 - t1)/2 * w_j * |x'(t)| (compute_intensities, :425-435).
  for (int p = 0; p < npatches; p++)
    if (!(lims[p + 1] > lims[p])) {
      err = "patch limits must be increasing";
      return false;
    }
#pragma omp parallel for schedule(static)
  for (int p = 0; p < npatches; p++) {
    const double a = lims[p], b = lims[p + 1];
    g.t1[p] = a;
    g.t2[p] = b;
    for (int j = 0; j < Np; j++) {
      const long long i = (long long)p * Np + j;
      const double t = ((b - a) / 2.0) * (g.fejer_nodes[j] + 1.0) + a;
      const double n0 = c.yp(t), n1 = -c.xp(t);
      const double nn = std::sqrt(n0 * n0 + n1 * n1);
      g.t[i] = t;
      g.x[i] = c.x(t);
      g.y[i] = c.y(t);
      g.nx[i] = n0 / nn;
      g.ny[i] = n1 / nn;
      g.w[i] = ((b - a) / 2.0) * g.fejer_weights[j] * nn;
    }
  }
  // Near-singular points: points of the two adjacent patches closer than delta
  // to the patch (determine_patch_near_singular_points, ForwardMap.tpp:224-285,
  // with the neighbour estimate of :164-218 for one closed curve). The
  // closest parameter is t1 if |x(t1) - x_q| < delta, else t2 if that holds,
  // else the interior minimiser of the distance (golden-section search).
  g.ns_idx.assign(npatches, {});
  g.ns_t.assign(npatches, {});
  if (ns_search) return ns_search(c, delta, g, err);
#pragma omp parallel for schedule(static)
  for (int p = 0; p < npatches; p++) {
    for (int e = 0; e < 2; e++) {
      const int pe = near_singular_candidate_patch(p, e, npatches);
      for (int j = 0; j < Np; j++) {
        const long long q = (long long)pe * Np + j;
        double ts;
        if (!near_singular_param(c, g.t1[p], g.t2[p], g.x[q], g.y[q], delta, ts)) continue;
        g.ns_idx[p].push_back(q);
        g.ns_t[p].push_back(ts);
      }
    }
  }
  return true;
}

// ------------------------------------------------------------ correction ---
void rp_rule(std::vector<double>& fn, std::vector<double>& fw) { fejer(40, fn, fw); }

void rp_columns(const Geometry& g, std::vector<RpColumn>& cols, std::vector<long long>& patch_col_off) {
  const int Np = g.Np;
  cols.clear();
  patch_col_off.assign(g.npatches + 1, 0);
  for (int pi = 0; pi < g.npatches; pi++) {
    patch_col_off[pi] = (long long)cols.size();
    for (int j = 0; j < Np; j++) {
      const long long i = (long long)pi * Np + j;
      cols.push_back(RpColumn{pi, g.t[i], g.t[i]});
    }
    for (size_t j = 0; j < g.ns_idx[pi].size(); j++)
      cols.push_back(RpColumn{pi, g.ns_t[pi][j], g.t[g.ns_idx[pi][j]]});
  }
  patch_col_off[g.npatches] = (long long)cols.size();
}

std::vector<double> rp_fold_table(const Geometry& g) {
  const int Np = g.Np;
  std::vector<double> T(Np * Np), F(Np * Np);
  for (int kk = 0; kk < Np; kk++) {
    const double x = g.fejer_nodes[kk];
    T[0 * Np + kk] = 1.0;
    T[1 * Np + kk] = x;
    for (int n = 2; n < Np; n++) T[n * Np + kk] = 2.0 * x * T[(n - 1) * Np + kk] - T[(n - 2) * Np + kk];
  }
  for (int n = 0; n < Np; n++)
    for (int kk = 0; kk < Np; kk++) F[n * Np + kk] = (n == 0 ? 1.0 : 2.0) / Np * T[n * Np + kk];
  return F;
}

void assemble_correction(const Geometry& g, const std::vector<zd>& W, const std::vector<long long>& patch_col_off,
                         Correction& out) {
  const int Np = g.Np;
  // target-major CSR: own patch first, then near-singular patches in patch order
  std::vector<std::vector<std::pair<int, long long>>> rows(g.N);  // (patch, global column)
  for (long long i = 0; i < g.N; i++) rows[i].push_back({(int)(i / Np), patch_col_off[i / Np] + i % Np});
  for (int pi = 0; pi < g.npatches; pi++)
    for (size_t j = 0; j < g.ns_idx[pi].size(); j++)
      rows[g.ns_idx[pi][j]].push_back({pi, patch_col_off[pi] + Np + (long long)j});
  out.ptr.assign(g.N + 1, 0);
  for (long long i = 0; i < g.N; i++) out.ptr[i + 1] = out.ptr[i] + (long long)rows[i].size();
  out.patch.resize(out.ptr[g.N]);
  out.w.resize((size_t)out.ptr[g.N] * Np);
  for (long long i = 0; i < g.N; i++)
    for (size_t e = 0; e < rows[i].size(); e++) {
      const long long ee = out.ptr[i] + (long long)e;
      out.patch[ee] = rows[i][e].first;
      std::copy(W.begin() + (size_t)rows[i][e].second * Np, W.begin() + (size_t)(rows[i][e].second + 1) * Np,
                out.w.begin() + (size_t)ee * Np);
    }
}

bool build_correction(const Geometry& g, const Curve& c, int F, double k, double eta, double p,
                      Correction& out, std::string& err) {
  const int Np = g.Np;
  std::vector<double> fn, fw;
  rp_rule(fn, fw);
  const std::vector<double> Tf = rp_fold_table(g);
  std::vector<RpColumn> cols;
  std::vector<long long> off;
  rp_columns(g, cols, off);
  // folded columns W[col][k] = sum_n Tf[n][k] M[col][n]
  std::vector<zd> W(cols.size() * Np);
#pragma omp parallel
  {
    std::vector<zd> M(Np);
#pragma omp for schedule(dynamic, 64)
    for (long long col = 0; col < (long long)cols.size(); col++) {
      const RpColumn& rc = cols[col];
      rp_column(c, F, k, eta, p, g.t1[rc.patch], g.t2[rc.patch], rc.tsing, rc.ttgt, Np, fn.data(), fw.data(),
                (int)fn.size(), M.data());
      for (int kk = 0; kk < Np; kk++) {
        zd s{0.0, 0.0};
        for (int n = 0; n < Np; n++) s = cadd(s, cscale(M[n], Tf[n * Np + kk]));
        W[(size_t)col * Np + kk] = s;
      }
    }
  }
  assemble_correction(g, W, off, out);
  (void)err;
  return true;
}

// ------------------------------------------------------------------ tree ---
static long long morton(long long px, long long py, int level) { return morton_code(px, py, level); }

namespace {
struct LevelGeom {
  double min_x, min_y, bs, h;
  long long nce, nrad, nbps;
  std::vector<double> radial;
  // InitializeConeSegments<Helmholtz>, Level.h:88-142
  LevelGeom(int level, double mx, double my, double bs_, double k) : min_x(mx), min_y(my), bs(bs_) {
    nbps = 1ll << level;
    const double minsize = (3.0 / 2.0 * bs) * (1.0 - 1e-10);
    const double maxsize = (std::sqrt(2.0) * bs * nbps - std::sqrt(2.0) / 2.0 * bs) * (1.0 + 1e-10);
    long long scal = 1;
    if (level > 1) scal = std::max<long long>(1, (long long)std::ceil(k * bs * M_1_PI * 1.2));
    nce = scal * 2;
    const long long nrad_on = scal;
    h = std::sqrt(2.0) / 2.0 * bs;
    const double eta = std::sqrt(2.0) / 3.0 - 1e-5;
    const double tmpr = (maxsize - minsize + 1e-10);
    const double tmps = std::min(eta / nrad_on, h * (1.0 / minsize - 1.0 / (minsize + tmpr)));
    const long long nact = (long long)std::ceil(h * (1.0 / minsize - 1.0 / maxsize) / tmps);
    radial.assign(nact + 1, 0.0);
    radial[0] = minsize;
    const double fs = (h * (1.0 / minsize - 1.0 / maxsize) / nact);
    for (long long i = nact - 1; i >= 0; i--) radial[nact - i] = h * maxsize / (h + maxsize * i * fs);
    nrad = nact;
  }
  long long point2morton(double x, double y, int level, bool& ok) const {
    const long long px = (long long)((x - min_x) / bs), py = (long long)((y - min_y) / bs);
    ok = !(px < 0 || py < 0 || px >= nbps || py >= nbps);
    return ok ? morton(px, py, level) : -1;
  }
  ConeGrid grid() const { return ConeGrid{min_x, min_y, bs, h, (int)nce, (int)nrad, radial.data()}; }
  long long nonrel(long long m, double x, double y) const { return grid_nonrel(grid(), m, x, y); }
  void cheb2cart(long long m, long long q, double s, double t, double& x, double& y) const {
    grid_cheb2cart(grid(), m, q, s, t, x, y);
  }
};
}  // namespace

bool build_tree(const double* x, const double* y, const double* nx, const double* ny, long long N,
                int nlev, double k, int Ps, int Pang, Tree& t, std::string& err, const std::vector<int>* perm_in,
                const double* tx, const double* ty, long long M, ConeBuilder cones, GroupBuilder groups) {
  if (nlev < 3) {
    err = "IFGF accelerator requires at least 3 levels.";
    return false;
  }
  if (N < 1 || N > (1ll << 30)) {
    err = "point count out of range";
    return false;
  }
  const bool sep = tx != nullptr;
  if (sep && (!ty || M < 1 || M > (1ll << 30) || perm_in)) {
    err = "target count out of range";
    return false;
  }
  t = Tree();
  t.nlev = nlev;
  t.D = nlev - 1;
  t.N = N;
  t.k = k;
  t.Ps = Ps;
  t.Pang = Pang;
  t.sep = sep;
  t.M = sep ? M : N;
  const int D = t.D, P = Ps * Pang;
  // Bounding box +- 1e-10 and the level-1 box size (BoxTree.h:179-231).
  double mn0 = 1e308, mn1 = 1e308, mx0 = -1e308, mx1 = -1e308;
  for (long long i = 0; i < N; i++) {
    mn0 = std::min(mn0, x[i]);
    mn1 = std::min(mn1, y[i]);
    mx0 = std::max(mx0, x[i]);
    mx1 = std::max(mx1, y[i]);
  }
  for (long long i = 0; sep && i < M; i++) {
    mn0 = std::min(mn0, tx[i]);
    mn1 = std::min(mn1, ty[i]);
    mx0 = std::max(mx0, tx[i]);
    mx1 = std::max(mx1, ty[i]);
  }
  if (!(mx0 - mn0 < 1e300) || !(mx1 - mn1 < 1e300)) {
    err = "non-finite point coordinates";
    return false;
  }
  mx0 += 1e-10;
  mx1 += 1e-10;
  mn0 -= 1e-10;
  mn1 -= 1e-10;
  const double bs0 = std::max(mx0 - mn0, mx1 - mn1);
  std::vector<LevelGeom> G;
  G.reserve(D + 1);
  for (int l = 0; l <= D; l++) G.emplace_back(std::max(l, 1), mn0, mn1, bs0 / (1ll << l), k);
  // SortBox (BoxTree.h:73-106): stable sort by leaf Morton code.
  std::vector<long long> code(N);
  for (long long i = 0; i < N; i++) {
    bool ok;
    code[i] = G[D].point2morton(x[i], y[i], D, ok);
    if (!ok) {
      err = "The 2D box index cannot be < 0 or larger than the number of boxes";
      return false;
    }
  }
  t.perm.resize(N);
  if (perm_in) {
    if ((long long)perm_in->size() != N) {
      err = "device Morton permutation has the wrong length";
      return false;
    }
    std::vector<char> seen(N, 0);
    for (long long i = 0; i < N; i++) {
      const int o = (*perm_in)[i];
      const bool ordered = i == 0 || code[(*perm_in)[i - 1]] < code[o] ||
                           (code[(*perm_in)[i - 1]] == code[o] && (*perm_in)[i - 1] < o);
      if (o < 0 || o >= N || seen[o] || !ordered) {
        err = "device Morton permutation differs from the stable sort of the leaf codes";
        return false;
      }
      seen[o] = 1;
    }
    t.perm = *perm_in;
  } else {
    std::iota(t.perm.begin(), t.perm.end(), 0);
    std::stable_sort(t.perm.begin(), t.perm.end(), [&](int a, int b) { return code[a] < code[b]; });
  }
  t.inv.resize(N);
  t.sx.resize(N);
  t.sy.resize(N);
  t.snx.resize(N);
  t.sny.resize(N);
  for (long long i = 0; i < N; i++) {
    const int o = t.perm[i];
    t.inv[o] = (int)i;
    t.sx[i] = x[o];
    t.sy[i] = y[o];
    t.snx[i] = nx[o];
    t.sny[i] = ny[o];
  }
  if (sep) {
    // targets: the same stable sort by leaf Morton code
    std::vector<long long> tcode(M);
    for (long long i = 0; i < M; i++) {
      bool ok;
      tcode[i] = G[D].point2morton(tx[i], ty[i], D, ok);
      if (!ok) {
        err = "The 2D box index cannot be < 0 or larger than the number of boxes";
        return false;
      }
    }
    t.tperm.resize(M);
    std::iota(t.tperm.begin(), t.tperm.end(), 0);
    std::stable_sort(t.tperm.begin(), t.tperm.end(), [&](int a, int b) { return tcode[a] < tcode[b]; });
    t.tinv.resize(M);
    t.tx.resize(M);
    t.ty.resize(M);
    for (long long i = 0; i < M; i++) {
      const int o = t.tperm[i];
      t.tinv[o] = (int)i;
      t.tx[i] = tx[o];
      t.ty[i] = ty[o];
    }
  }
  // Relevant boxes per level (BoxTree.h:111-174): leaf boxes from the sorted
  // codes, parents = codes / 4, each level ascending.
  std::vector<std::vector<long long>> boxes(D + 1);
  t.leaf_of_point.resize(N);
  t.leaf_begin.clear();
  for (long long i = 0; i < N; i++) {
    const long long c = code[t.perm[i]];
    if (boxes[D].empty() || boxes[D].back() != c) {
      boxes[D].push_back(c);
      t.leaf_begin.push_back((int)i);
    }
    t.leaf_of_point[i] = (int)boxes[D].size() - 1;
  }
  t.leaf_begin.push_back((int)N);
  for (int l = D - 1; l >= 2; l--)
    for (long long c : boxes[l + 1])
      if (boxes[l].empty() || boxes[l].back() != c / 4) boxes[l].push_back(c / 4);
  std::vector<std::unordered_map<long long, int>> box_index(D + 1);
  t.lv.assign(D + 1, LevelInfo());
  int box_off = 0;
  for (int l = 2; l <= D; l++) {
    for (size_t b = 0; b < boxes[l].size(); b++) box_index[l][boxes[l][b]] = (int)b;
    LevelInfo& L = t.lv[l];
    L.level = l;
    L.min_x = mn0;
    L.min_y = mn1;
    L.bs = G[l].bs;
    L.h = G[l].h;
    L.nce = (int)G[l].nce;
    L.nrad = (int)G[l].nrad;
    L.radial_off = (int)t.radial.size();
    t.radial.insert(t.radial.end(), G[l].radial.begin(), G[l].radial.end());
    L.nbox = (int)boxes[l].size();
    L.box_off = box_off;
    box_off += L.nbox;
    L.nslot = (int)((2 * G[l].nce + 1) * G[l].nrad);
    for (long long m : boxes[l]) {
      double cx, cy;
      grid_center(G[l].grid(), m, cx, cy);
      t.box_morton.push_back(m);
      t.box_cx.push_back(cx);
      t.box_cy.push_back(cy);
    }
  }
  // children slots (GetMortonRelChildren, BoxTree.h:279-297)
  t.box_child.assign((size_t)box_off * 4, -1);
  for (int l = 2; l < D; l++)
    for (int b = 0; b < t.lv[l].nbox; b++)
      for (int cp = 0; cp < 4; cp++) {
        auto it = box_index[l + 1].find(boxes[l][b] * 4 + cp);
        if (it != box_index[l + 1].end()) t.box_child[(size_t)(t.lv[l].box_off + b) * 4 + cp] = it->second;
      }
  // relevant (source) boxes of the 3x3 neighbourhood of the level-l box with
  // Morton code m (GetNeighbours, Level.h:378-404) and its cousins (children
  // of the parent's neighbours that are not neighbours, GetCousins,
  // Level.h:261-289), as local box indices in ascending Morton order
  auto neighbours_of = [&](int l, long long m, std::vector<int>& out) {
    const long long bx = (long long)morton_merge2((uint64_t)m);
    const long long by = (long long)morton_merge2((uint64_t)m >> 1);
    std::vector<std::pair<long long, int>> nb;
    for (long long j = std::max(0ll, by - 1); j <= std::min(G[l].nbps - 1, by + 1); j++)
      for (long long i = std::max(0ll, bx - 1); i <= std::min(G[l].nbps - 1, bx + 1); i++) {
        auto it = box_index[l].find(morton(i, j, l));
        if (it != box_index[l].end()) nb.push_back({it->first, it->second});
      }
    std::sort(nb.begin(), nb.end());
    out.clear();
    for (auto& e : nb) out.push_back(e.second);
  };
  auto cousins_of = [&](int l, long long m, std::vector<int>& out) {
    const long long bx = (long long)morton_merge2((uint64_t)m);
    const long long by = (long long)morton_merge2((uint64_t)m >> 1);
    const long long xlo = std::max<long long>((bx / 2) * 2 - 2, 0), xup = std::min<long long>(G[l].nbps - 1, (bx / 2) * 2 + 3);
    const long long ylo = std::max<long long>((by / 2) * 2 - 2, 0), yup = std::min<long long>(G[l].nbps - 1, (by / 2) * 2 + 3);
    std::vector<std::pair<long long, int>> c;
    for (long long j = ylo; j <= yup; j++)
      for (long long i = xlo; i <= xup; i++) {
        if (std::abs(i - bx) <= 1 && std::abs(j - by) <= 1) continue;
        auto it = box_index[l].find(morton(i, j, l));
        if (it != box_index[l].end()) c.push_back({it->first, it->second});
      }
    std::sort(c.begin(), c.end());
    out.clear();
    for (auto& e : c) out.push_back(e.second);
  };
  if (groups && !cones) {
    err = "the grouping step on the device needs the device cone builder";
    return false;
  }
  // leaf neighbours of every source leaf box
  if (!groups) {
    t.nb_ptr.assign(1, 0);
    std::vector<int> nb;
    for (size_t b = 0; b < boxes[D].size(); b++) {
      neighbours_of(D, boxes[D][b], nb);
      t.nb_list.insert(t.nb_list.end(), nb.begin(), nb.end());
      t.nb_ptr.push_back((int)t.nb_list.size());
    }
  }
  // Target groups per level: targets sharing their level-l box. For the
  // forward map these are the relevant boxes themselves (every target is a
  // source); for separate targets, the runs of equal level-l codes of the
  // sorted targets (codes are ascending at every level). cz[l][grp] is the
  // group's cousin list, pbox[i * (D+1) + l] the group of sorted target i.
  const long long NT = t.M;
  const double* TX = sep ? t.tx.data() : t.sx.data();
  const double* TY = sep ? t.ty.data() : t.sy.data();
  std::vector<std::vector<std::vector<int>>> cz(D + 1);
  std::vector<int> pbox;
  ConeInput cin;
  bool bad = false;
  if (groups) {
    GroupJob job;
    job.tx = TX;
    job.ty = TY;
    job.nt = NT;
    job.sep = sep;
    job.nbps.assign(D + 1, 0);
    for (int l = 0; l <= D; l++) job.nbps[l] = G[l].nbps;
    if (!groups(job, t, pbox, cin, err)) return false;
    for (long long i = 0; sep && i < M; i++) {
      const int g = t.tleaf[i];
      for (int n = t.tnb_ptr[g]; n < t.tnb_ptr[g + 1]; n++)
        t.n_near_t += t.leaf_begin[t.tnb_list[n] + 1] - t.leaf_begin[t.tnb_list[n]];
    }
  } else if (!sep) {
    pbox.assign((size_t)NT * (D + 1), -1);
    for (int l = 2; l <= D; l++) {
      cz[l].resize(boxes[l].size());
      for (size_t b = 0; b < boxes[l].size(); b++) cousins_of(l, boxes[l][b], cz[l][b]);
    }
    // per sorted point: box at each level (Point2Morton at that level)
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (long long i = 0; i < N; i++)
      for (int l = 2; l <= D; l++) {
        bool ok;
        const long long m = G[l].point2morton(t.sx[i], t.sy[i], l, ok);
        auto it = ok ? box_index[l].find(m) : box_index[l].end();
        if (it == box_index[l].end()) {
          bad = true;
          continue;
        }
        pbox[(size_t)i * (D + 1) + l] = it->second;
      }
    if (bad) {
      err = "point outside the relevant boxes of a level";
      return false;
    }
  } else {
    pbox.assign((size_t)NT * (D + 1), -1);
    std::vector<long long> gcode;
    for (int l = 2; l <= D; l++) {
      gcode.clear();
      for (long long i = 0; i < M; i++) {
        bool ok;
        const long long m = G[l].point2morton(TX[i], TY[i], l, ok);
        if (!ok) {
          err = "The 2D box index cannot be < 0 or larger than the number of boxes";
          return false;
        }
        if (gcode.empty() || gcode.back() != m) gcode.push_back(m);
        pbox[(size_t)i * (D + 1) + l] = (int)gcode.size() - 1;
      }
      cz[l].resize(gcode.size());
      for (size_t gi = 0; gi < gcode.size(); gi++) cousins_of(l, gcode[gi], cz[l][gi]);
      if (l == D) {
        // near field of the targets: source leaf boxes around each leaf group
        t.tnb_ptr.assign(1, 0);
        std::vector<int> nb;
        for (size_t gi = 0; gi < gcode.size(); gi++) {
          neighbours_of(D, gcode[gi], nb);
          t.tnb_list.insert(t.tnb_list.end(), nb.begin(), nb.end());
          t.tnb_ptr.push_back((int)t.tnb_list.size());
        }
      }
    }
    t.tleaf.resize(M);
    for (long long i = 0; i < M; i++) {
      const int g = pbox[(size_t)i * (D + 1) + D];
      t.tleaf[i] = g;
      for (int n = t.tnb_ptr[g]; n < t.tnb_ptr[g + 1]; n++)
        t.n_near_t += t.leaf_begin[t.tnb_list[n] + 1] - t.leaf_begin[t.tnb_list[n]];
    }
  }
  // interpolation entries: per target, levels D..2, cousins ascending
  // (Interpolation, BoxTree.h:876-937); first the non-relevant cone ids.
  if (!groups) {
    t.it_ptr.assign(NT + 1, 0);
    for (long long i = 0; i < NT; i++) {
      long long c = 0;
      for (int l = D; l >= 2; l--) c += (long long)cz[l][pbox[(size_t)i * (D + 1) + l]].size();
      t.it_ptr[i + 1] = t.it_ptr[i] + c;
    }
  }
  const long long nent = t.it_ptr[NT];
  if (nent > 2147483647LL) {
    err = "too many interpolation entries";
    return false;
  }
  // Dense (box, non-relevant id) maps and protoboxes, level by level;
  // interpolation points (Interpolator.h:37-41): s = xs[j % Ps], theta = xang[j / Ps].
  long long map_total = 0, proto_total = 0;
  for (int l = 2; l <= D; l++) {
    t.lv[l].map_off = map_total;
    map_total += (long long)t.lv[l].nbox * t.lv[l].nslot;
    if (l > 2) {
      t.lv[l].proto_off = proto_total;
      proto_total += (long long)t.lv[l - 1].nslot * 4 * P;
    }
  }
  std::vector<double> xs(Ps), xa(Pang);
  for (int i = 0; i < Ps; i++) xs[Ps - 1 - i] = std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * Ps));
  for (int i = 0; i < Pang; i++) xa[Pang - 1 - i] = std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * Pang));
  if (cones) {
    cin.tx = TX;
    cin.ty = TY;
    cin.nt = NT;
    cin.pbox = pbox.data();
    cin.xs = xs;
    cin.xa = xa;
    if (!groups) {
      cin.grp_off.assign(D + 2, 0);
      long long ngrp = 0, ncz = 0;
      for (int l = 2; l <= D; l++) {
        cin.grp_off[l] = ngrp;
        ngrp += (long long)cz[l].size();
        for (const auto& v : cz[l]) ncz += (long long)v.size();
      }
      cin.grp_off[D + 1] = ngrp;
      if (ncz > 2147483647LL) {
        err = "too many cousin entries";
        return false;
      }
      cin.cz_ptr.reserve(ngrp + 1);
      cin.cz_list.reserve(ncz);
      cin.cz_ptr.push_back(0);
      for (int l = 2; l <= D; l++)
        for (const auto& v : cz[l]) {
          cin.cz_list.insert(cin.cz_list.end(), v.begin(), v.end());
          cin.cz_ptr.push_back((int)cin.cz_list.size());
        }
    }
    if (!cones(cin, t, err)) return false;
  } else {
    std::vector<int> ent_q(nent);
#pragma omp parallel for schedule(dynamic, 256) reduction(|| : bad)
    for (long long i = 0; i < NT; i++) {
      long long e = t.it_ptr[i];
      for (int l = D; l >= 2; l--)
        for (int c : cz[l][pbox[(size_t)i * (D + 1) + l]]) {
          const long long q = G[l].nonrel(boxes[l][c], TX[i], TY[i]);
          if (q < 0) bad = true;
          ent_q[e++] = (int)q;
        }
    }
    if (bad) {
      err = "target outside the cone grid of a cousin box";
      return false;
    }
    // mark relevant cones, level by level top-down, then number them (box
    // ascending, non-relevant id ascending) -- InitializeRelevantConeSegments,
    // BoxTree.h:553-583 with the merge of :430-459 and flattening of :461-514.
    t.dense_map.assign(map_total, -1);
    t.proto.assign(proto_total, -1);
    std::vector<char> mark(map_total, 0);
    for (long long i = 0; i < NT; i++) {
      long long e = t.it_ptr[i];
      for (int l = D; l >= 2; l--)
        for (int c : cz[l][pbox[(size_t)i * (D + 1) + l]]) mark[t.lv[l].map_off + (long long)c * t.lv[l].nslot + ent_q[e++]] = 1;
    }
    int cone_off = 0;
    for (int l = 2; l <= D; l++) {
      LevelInfo& L = t.lv[l];
      if (l > 2) {
        // cones needed by the parents (GetRelevantConeSegmentsDueToParents,
        // BoxTree.h:299-428): protobox[q][cp] is filled from the first parent
        // cone with non-relevant id q that has a child at position cp.
        const LevelInfo& Pl = t.lv[l - 1];
        int* proto = &t.proto[L.proto_off];
        for (int g = Pl.cone_off; g < Pl.cone_off + Pl.ncone; g++) {
          const int pb = t.cone_box[g];
          const long long q = (long long)t.cone_post[g] * Pl.nrad + t.cone_posr[g];
          for (int cp = 0; cp < 4; cp++) {
            const int cb = t.box_child[(size_t)(Pl.box_off + pb) * 4 + cp];
            if (cb < 0) continue;
            int* slot = proto + ((size_t)q * 4 + cp) * P;
            if (slot[0] < 0) {
              for (int j = 0; j < P; j++) {
                double px, py;
                G[l - 1].cheb2cart(boxes[l - 1][pb], q, xs[j % Ps], xa[j / Ps], px, py);
                const long long cq = G[l].nonrel(boxes[l][cb], px, py);
                if (cq < 0) {
                  err = "parent interpolation node outside the child cone grid";
                  return false;
                }
                slot[j] = (int)cq;
              }
            }
            for (int j = 0; j < P; j++) mark[L.map_off + (long long)cb * L.nslot + slot[j]] = 1;
            t.n_prop += P;
          }
        }
      }
      L.cone_off = cone_off;
      int nc = 0;
      for (int b = 0; b < L.nbox; b++)
        for (int q = 0; q < L.nslot; q++) {
          const long long a = L.map_off + (long long)b * L.nslot + q;
          if (!mark[a]) continue;
          t.dense_map[a] = nc++;
          t.cone_box.push_back(b);
          t.cone_posr.push_back(q % L.nrad);
          t.cone_post.push_back(q / L.nrad);
          t.cone_level.push_back((unsigned char)l);
        }
      if (nc == 0 && !sep) {  // a field's coarse levels may have none (its targets are near every source)
        err = "The number of cone segments on any level cannot be 0";
        return false;
      }
      L.ncone = nc;
      cone_off += nc;
    }
    t.n_cones = cone_off;
    // final interpolation entries as global cone ids
    t.it_cone.resize(nent);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < NT; i++) {
      long long e = t.it_ptr[i];
      for (int l = D; l >= 2; l--)
        for (int c : cz[l][pbox[(size_t)i * (D + 1) + l]]) {
          t.it_cone[e] = t.lv[l].cone_off + t.dense_map[t.lv[l].map_off + (long long)c * t.lv[l].nslot + ent_q[e]];
          e++;
        }
    }
  }
  for (int g = t.lv[D].cone_off; g < t.lv[D].cone_off + t.lv[D].ncone; g++)
    t.n_leaf += (long long)P * (t.leaf_begin[t.cone_box[g] + 1] - t.leaf_begin[t.cone_box[g]]);
  t.n_interp = nent;
  if (!sep) t.point_box.swap(pbox);
  return true;
}


bool build_exclusions(const Geometry& g, const Tree& t, std::vector<int>& patch_of_sorted,
                      std::vector<int>& excl, long long& n_near, std::string& err) {
  const long long N = g.N;
  const int Np = g.Np;
  // own patch, then the patches the point is near-singular to, in patch order
  // (patch order index i); exn[i] entries in ex[i * kMaxExcl ..]
  std::vector<int> ex((size_t)N * kMaxExcl, -1), exn(N, 1);
  for (long long i = 0; i < N; i++) ex[(size_t)i * kMaxExcl] = (int)(i / Np);
  for (int p = 0; p < g.npatches; p++)
    for (long long q : g.ns_idx[p]) {
      int* e = &ex[(size_t)q * kMaxExcl];
      if (std::find(e, e + exn[q], p) != e + exn[q]) continue;
      if (exn[q] == kMaxExcl) {
        err = "a point is near-singular to more patches than the near-field exclusion list holds";
        return false;
      }
      e[exn[q]++] = p;
    }
  patch_of_sorted.resize(N);
  excl.assign((size_t)N * kMaxExcl, -1);
  for (long long i = 0; i < N; i++) {
    const int s = t.inv[i];
    patch_of_sorted[s] = (int)(i / Np);
    std::copy(&ex[(size_t)i * kMaxExcl], &ex[(size_t)i * kMaxExcl] + exn[i], &excl[(size_t)s * kMaxExcl]);
  }
  // every excluded pair must be a leaf-level neighbour pair
  // (CheckIfSingandNearSingInNeighborhood, BoxTree.h:1079-1161)
  const int D = t.D;
  auto leaf_xy = [&](long long sorted, long long& bx, long long& by) {
    const long long m = t.box_morton[t.lv[D].box_off + t.leaf_of_point[sorted]];
    bx = (long long)morton_merge2((uint64_t)m);
    by = (long long)morton_merge2((uint64_t)m >> 1);
  };
  bool far = false;
#pragma omp parallel for schedule(static) reduction(|| : far)
  for (long long i = 0; i < N; i++) {
    long long ax, ay;
    leaf_xy(t.inv[i], ax, ay);
    for (int e = 0; e < exn[i]; e++) {
      const int p = ex[(size_t)i * kMaxExcl + e];
      for (long long j = (long long)p * Np; j < (long long)(p + 1) * Np; j++) {
        long long bx, by;
        leaf_xy(t.inv[j], bx, by);
        if (std::abs(ax - bx) > 1 || std::abs(ay - by) > 1) far = true;
      }
    }
  }
  if (far) {
    err = "Singular or near singular point not in the leaf neighbourhood (use fewer levels)";
    return false;
  }
  // near-field pair count (SingularInteractions, BoxTree.h:614-683)
  long long cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
  for (long long i = 0; i < N; i++) {
    const int b = t.leaf_of_point[i];
    const int* e = &excl[(size_t)i * kMaxExcl];
    for (int n = t.nb_ptr[b]; n < t.nb_ptr[b + 1]; n++) {
      const int nb = t.nb_list[n];
      for (int s = t.leaf_begin[nb]; s < t.leaf_begin[nb + 1]; s++) {
        const int ps = patch_of_sorted[s];
        if (ps == e[0] || ps == e[1] || ps == e[2] || ps == e[3]) continue;
        cnt++;
      }
    }
  }
  n_near = cnt;
  return true;
}

}  // namespace ifgf
