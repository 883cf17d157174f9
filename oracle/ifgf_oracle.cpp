// ============================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference IFGF2D forward map (manuelarturosantana/
// IFGF2D), written from the reference's algorithm, used ONLY by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, and only as the
// checker. It is never linked into the product library (ifgf2d_amd/).
//
// Hankel functions come from the reference's own Amos library (ZBESH) compiled
// from complex_bessel-master/src into oracle/_ref/libamos_ref.so by
// oracle/Makefile -- exactly what the reference calls per kernel evaluation
// (besselFunctions.h:312-341 -> amos_iso_c_fortran_wrapper.f90:113-135).
//
// Every routine cites the reference file:line it follows. Deviations:
//  * DoubleLayer / Combined kernels follow the mathematics (the reference's are
//    commented out or buggy: GreenFunctions.hpp:55-57,68-75); x = target,
//    y = source, n = n_y throughout.
//  * Relevant-box iteration order is ascending Morton code (the reference
//    iterates a std::unordered_map, BoxTree.h:468-487), so every relevant-cone
//    list and every "first wins" protobox entry is reproducible.
//  * The bounding-box pre-filter of the near-singular search uses exact patch
//    extrema from a sampled sign-change search instead of the colleague-matrix
//    eigenvalues (Chebyshev1d.hpp:231-267); it only filters, the decision is
//    the distance test of ForwardMap.tpp:253-275.
//  * In the RP precomputation the separation x - y between the target and a
//    quadrature node is formed from the parameter difference with
//    trigonometric difference identities (Curve::diff), not by subtracting
//    Cartesian coordinates (ForwardMap.tpp:355-365): the double-layer
//    numerator (x - y).n_y cancels catastrophically otherwise (the RP nodes
//    come within ~1e-10 of the singular point), which made the folded
//    correction sensitive to the last bit of the point coordinates.
//  * The interior minimiser of that distance test is golden-section search to
//    1e-12 instead of Boost's Brent (BrentsMethod.hpp:42-156); for closed
//    curves split into small patches the minimum lies at an endpoint and this
//    branch does not decide anything.
// Pinned by tests/test_oracle.py and tests/test_geometry.py (mpmath golden
// vectors, the reference's known answers, circle eigenvalues, the Mie series,
// curve derivatives and integrals against mpmath, an interior point-source
// closed form on the kite and stars). Parity unpinned: the exact iterates of
// Eigen's Householder GMRES (Eigen is absent; only its residual-minimising
// property and a dense solve are checked) and Boost's Brent minimiser.
// OpenMP: the per-target / per-cone loops of an application run in parallel;
// every output is summed in the serial order, so the results do not depend
// on the thread count.
// ============================================================================
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

using cd = std::complex<double>;
using ll = long long;

extern "C" void zbesh_wrap(double zr, double zi, double order, int kode, int M, int N, double* cyr,
                           double* cyi, int* nz, int* ierr);

namespace orc {

static thread_local std::string g_err;

// ----------------------------------------------------------------------------
// Hankel function H_n^(1)(x), x real > 0: besselFunctions.h:312-341.
cd hankel1(int n, double x) {
  double cr = 0, ci = 0;
  int nz = 0, ierr = 0;
  zbesh_wrap(x, 0.0, (double)n, 1, 1, 1, &cr, &ci, &nz, &ierr);
  return cd(cr, ci);
}

enum Formulation { kSingleLayer = 0, kDoubleLayer = 1, kCombined = 2 };

// GF<F>, GreenFunctions.hpp:86-109 (+ GF_far :35-83). tx,ty = target, sx,sy =
// source, nx,ny = unit source normal. SingleLayer = (i/4) H0(k r) as the
// reference; DoubleLayer = dG/dn_y = (ik/4) H1(kr) (x-y).n_y / r; Combined =
// DoubleLayer - i*eta*SingleLayer.
cd kernel_d(int F, double dx, double dy, double nx, double ny, double k, double eta) {
  const double r = std::sqrt(dx * dx + dy * dy);
  if (r < 1e-14) return cd(0.0, 0.0);  // GreenFunctions.hpp:97-100
  const cd I(0.0, 1.0);
  if (F == kSingleLayer) return I / 4.0 * hankel1(0, k * r);
  const cd dl = I * k / 4.0 * hankel1(1, k * r) * ((dx * nx + dy * ny) / r);
  if (F == kDoubleLayer) return dl;
  return dl - I * eta * (I / 4.0 * hankel1(0, k * r));
}
cd kernel(int F, double tx, double ty, double sx, double sy, double nx, double ny, double k,
          double eta) {
  return kernel_d(F, tx - sx, ty - sy, nx, ny, k, eta);
}

// factorization(), GreenFunctions.hpp:125-132: exp(i k r)/sqrt(r).
cd factorization(double r, double k) {
  return std::exp(cd(0.0, 1.0) * k * r) / std::sqrt(r);
}

// ----------------------------------------------------------------------------
// Quadrature.hpp:13-65
double ab2cd(double x, double a, double b, double c, double d) {
  return ((d - c) / (b - a)) * (x - a) + c;
}
double ab2cdjac(double a, double b, double c, double d) { return (d - c) / (b - a); }

void fejer1(int n, std::vector<double>& nodes, std::vector<double>& weights) {
  nodes.assign(n, 0.0);
  weights.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    nodes[i] = std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * n));
    double w = 0.0;
    for (int j = 1; j <= (int)std::floor(n * 0.5); j++)
      w += std::cos(j * M_PI * (2.0 * i + 1.0) / n) / (4.0 * j * j - 1.0);
    weights[i] = (2.0 / n) * (1.0 - 2.0 * w);
  }
}

// RP changes of variables, Quadrature.hpp:73-147.
double v_func(double t, double p) {
  double val = (0.5 - (1.0 / p)) * std::pow((t / M_PI) - 1.0, 3);
  return val + (1.0 / p) * ((t / M_PI) - 1.0) + 0.5;
}
double vp_func(double t, double p) {
  return (3.0 / M_PI) * (0.5 - (1.0 / p)) * std::pow((t / M_PI) - 1.0, 2.0) + 1.0 / (p * M_PI);
}
double w_cov_base(double t, double p) {
  double vp = std::pow(v_func(t, p), p);
  return 2.0 * M_PI * vp / (vp + std::pow(v_func(2 * M_PI - t, p), p));
}
double wp_cov_base(double t, double p) {
  double val = 2.0 * M_PI * p * vp_func(t, p);
  val *= std::pow(v_func(t, p), p - 1.0) * std::pow(v_func(2 * M_PI - t, p), p - 1.0);
  return val / std::pow(std::pow(v_func(t, p), p) + std::pow(v_func(2 * M_PI - t, p), p), 2.0);
}
int sgn(double v) { return (0.0 < v) - (v < 0.0); }
double xi_cov_center(double t, double a, double p) {
  return a + ((sgn(t) - a) / M_PI) * w_cov_base(M_PI * std::abs(t), p);
}
double xi_cov_right(double t, double p) {
  return 1.0 - ((2.0 / M_PI) * w_cov_base(M_PI * std::abs((t - 1.0) / 2.0), p));
}
double xi_cov_left(double t, double p) {
  return -1.0 + ((2.0 / M_PI) * w_cov_base(M_PI * std::abs((t + 1.0) / 2.0), p));
}
double dxi_cov_center(double t, double a, double p) {
  return (1.0 - a * sgn(t)) * wp_cov_base(M_PI * std::abs(t), p);
}
double dxi_cov_right(double t, double p) {
  double tmp = (t - 1.0) / 2.0;
  return -1.0 * wp_cov_base(M_PI * std::abs(tmp), p) * sgn(tmp);
}
double dxi_cov_left(double t, double p) {
  double tmp = (t + 1.0) / 2.0;
  return wp_cov_base(M_PI * std::abs(tmp), p) * sgn(tmp);
}

// ----------------------------------------------------------------------------
// Curves, Geometry/Curve.hpp:21-161 and Curve.cpp:105-138. normal_t = (y', -x')
// (Curve.hpp:97-102). Star is new: r(t) = R(1 + sum_m a_m cos mt + b_m sin mt).
struct Curve {
  int type = 0;  // 0 circle, 1 kite, 2 star
  double cx = 0, cy = 0, R = 1;
  double A = 1, B = 0.65, C = -0.65, D = 1.5;
  std::vector<double> am, bm;  // star modes m = 0.. (index = m)
  double tlim1 = 0, tlim2 = 2 * M_PI;

  double star_r(double t, double& rp) const {
    double r = 1.0, d = 0.0;
    for (size_t m = 0; m < am.size(); m++) {
      r += am[m] * std::cos(m * t) + bm[m] * std::sin(m * t);
      d += -am[m] * m * std::sin(m * t) + bm[m] * m * std::cos(m * t);
    }
    rp = R * d;
    return R * r;
  }
  double xt(double t) const {
    if (type == 0) return R * std::cos(t) + cx;
    if (type == 1) return A * std::cos(t) + B * std::cos(2.0 * t) + C;
    double rp, r = star_r(t, rp);
    return cx + r * std::cos(t);
  }
  double yt(double t) const {
    if (type == 0) return R * std::sin(t) + cy;
    if (type == 1) return D * std::sin(t);
    double rp, r = star_r(t, rp);
    return cy + r * std::sin(t);
  }
  double xpt(double t) const {
    if (type == 0) return -R * std::sin(t);
    if (type == 1) return -A * std::sin(t) - 2.0 * B * std::sin(2.0 * t);
    double rp, r = star_r(t, rp);
    return rp * std::cos(t) - r * std::sin(t);
  }
  double ypt(double t) const {
    if (type == 0) return R * std::cos(t);
    if (type == 1) return D * std::cos(t);
    double rp, r = star_r(t, rp);
    return rp * std::sin(t) + r * std::cos(t);
  }
  // (x(a) - x(b), y(a) - y(b)) with b = a - d, from
  // cos a - cos b = -2 sin((a+b)/2) sin((a-b)/2), sin a - sin b = 2 cos((a+b)/2) sin((a-b)/2).
  void diff(double a, double d, double& dx, double& dy) const {
    const double b = a - d, h = 0.5 * (a + b), sd = std::sin(0.5 * d);
    if (type == 0) {
      dx = R * (-2.0 * std::sin(h) * sd);
      dy = R * (2.0 * std::cos(h) * sd);
    } else if (type == 1) {
      dx = A * (-2.0 * std::sin(h) * sd) + B * (-2.0 * std::sin(2.0 * h) * std::sin(d));
      dy = D * (2.0 * std::cos(h) * sd);
    } else {
      double rp, ra = star_r(a, rp), rb = star_r(b, rp);
      double dr = 0.0;
      for (size_t m = 0; m < am.size(); m++) {
        const double s = std::sin(0.5 * m * d);
        dr += am[m] * (-2.0 * std::sin(m * h) * s) + bm[m] * (2.0 * std::cos(m * h) * s);
      }
      dr *= R;
      (void)rb;
      dx = ra * (-2.0 * std::sin(h) * sd) + dr * std::cos(b);
      dy = ra * (2.0 * std::cos(h) * sd) + dr * std::sin(b);
    }
  }
};

// Curve::comp_segement_length / ref_seg_len / compute_patch_lims,
// Curve.cpp:3-69 (200-node Fejer rule, bisection until len <= size_lim).
double seg_length(const Curve& c, double a, double b) {
  std::vector<double> nd, wt;
  fejer1(200, nd, wt);
  double len = 0.0;
  for (int i = 0; i < 200; i++) {
    double t = ab2cd(nd[i], -1.0, 1.0, a, b);
    len += std::sqrt(c.xpt(t) * c.xpt(t) + c.ypt(t) * c.ypt(t)) * wt[i];
  }
  return len * ab2cdjac(-1.0, 1.0, a, b);
}
void ref_seg_len(const Curve& c, double a, double b, double lim, std::vector<double>& out) {
  if (seg_length(c, a, b) <= lim) {
    out.push_back(b);
    return;
  }
  double mid = (a + b) / 2.0;
  ref_seg_len(c, a, mid, lim, out);
  ref_seg_len(c, mid, b, lim, out);
}
std::vector<double> compute_patch_lims(const Curve& c, double wl_per_patch, double k) {
  double wl = 2 * M_PI / k;
  double len = seg_length(c, c.tlim1, c.tlim2);
  int nwl = (int)std::ceil(len / wl);
  int nint = (int)std::ceil(nwl / wl_per_patch);
  std::vector<double> lims{c.tlim1};
  double tl = c.tlim2 - c.tlim1;
  for (int i = 0; i < nint; i++)
    ref_seg_len(c, c.tlim1 + tl * i / (double)nint, c.tlim1 + tl * (i + 1) / (double)nint,
                wl * wl_per_patch, lims);
  return lims;
}

// ----------------------------------------------------------------------------
// Level, IFGF_Source/Level.h:18-455.
uint64_t splitBy2(unsigned int a, int level) {  // Level.h:190-206
  uint64_t x = a;
  x &= (static_cast<uint64_t>(1) << (level + 1)) - 1;
  x = (x | (x << 16)) & 0xFFFF0000FFFFull;
  x = (x | (x << 8)) & 0xFF00FF00FF00FFull;
  x = (x | (x << 4)) & 0xF0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
unsigned int mergeFrom2(uint64_t x) {  // Level.h:215-226
  x &= 0x5555555555555555ull;
  x = (x ^ (x >> 1)) & 0x3333333333333333ull;
  x = (x ^ (x >> 2)) & 0xF0F0F0F0F0F0F0Full;
  x = (x ^ (x >> 4)) & 0xFF00FF00FF00FFull;
  x = (x ^ (x >> 8)) & 0xFFFF0000FFFFull;
  x = (x ^ (x >> 16)) & 0xffffffffull;
  return (unsigned int)x;
}
ll box2morton(ll x, ll y, int level) {  // Level.h:172-181
  return (ll)(splitBy2((unsigned)x, level) | (splitBy2((unsigned)y, level) << 1));
}
ll point2morton(double x, double y, double mx, double my, double bs, int level) {  // :145-161
  const ll px = static_cast<ll>((x - mx) / bs);
  const ll py = static_cast<ll>((y - my) / bs);
  if (px < 0 || py < 0 || px >= (1ll << level) || py >= (1ll << level))
    throw std::logic_error("The 2D box index cannot be < 0 or larger than the number of boxes");
  return box2morton(px, py, level);
}
// Functions::Atan2 / CartToPol / PolToCart, PolarCoordinates.h:8-57.
double atan2_ref(double y, double x) {
  if (x > 0 && y >= 0) return std::atan(y / x);
  if (x > 0 && y < 0) return std::atan(y / x) + 2.0 * M_PI;
  if (x < 0 && y > 0) return std::atan(y / x) + M_PI;
  if (x < 0 && y == 0) return M_PI;
  if (x < 0 && y < 0) return std::atan(y / x) + M_PI;
  if (x == 0 && y > 0) return M_PI_2;
  if (x == 0 && y < 0) return 3.0 * M_PI_2;
  return 0;
}
void cart2pol(double& x, double& y) {
  const double r = std::sqrt(x * x + y * y);
  if (r == 0.0) {
    x = 0.0;
    y = 0.0;
    return;
  }
  const double th = atan2_ref(y, x);
  x = r;
  y = th;
}
void pol2cart(double& r, double& th) {
  double s = std::sin(th), c = std::cos(th);
  const double x = r * c, y = r * s;
  r = x;
  th = y;
}

struct Level {
  int level, nlevels;
  double min_x, min_y, boxsize;
  ll nbps;
  std::vector<ll> rel2box, rel2nonrel;
  double h;  // scaling_r2s_
  ll nce;    // nconeselevation_
  std::vector<double> radial;
  std::unordered_set<ll> relboxes;
  std::unordered_map<ll, std::array<ll, 2>> box2pts;
  std::unordered_map<ll, std::unordered_map<ll, ll>> boxnonrel2rel;
  // protobox_[parent nonrel][child pos] -> (child nonrel -> node ids), Level.h:70
  std::unordered_map<ll, std::array<std::map<ll, std::vector<int>>, 4>> protobox;

  Level(int lv, int nl, double mx, double my, double bs, double k)
      : level(lv), nlevels(nl), min_x(mx), min_y(my), boxsize(bs), nbps(1ll << lv) {
    // InitializeConeSegments<Helmholtz>, Level.h:88-142
    const double minsize = (3.0 / 2.0 * boxsize) * (1.0 - 1e-10);
    const double maxsize =
        (std::sqrt(2.0) * boxsize * nbps - std::sqrt(2.0) / 2.0 * boxsize) * (1.0 + 1e-10);
    ll scal = 1;
    if (level > 1) scal = std::max<ll>(1, (ll)std::ceil(k * boxsize * M_1_PI * 1.2));
    nce = scal * 2;
    const ll nrad = scal * 1;
    h = std::sqrt(2.0) / 2.0 * boxsize;
    const double eta = std::sqrt(2.0) / 3.0 - 1e-5;
    const double tmpr = (maxsize - minsize + 1e-10);
    const double tmps = std::min(eta / nrad, h * (1.0 / minsize - 1.0 / (minsize + tmpr)));
    const ll nact = (ll)std::ceil(h * (1.0 / minsize - 1.0 / maxsize) / tmps);
    radial.assign(nact + 1, 0.0);
    radial[0] = minsize;
    const double fs = (h * (1.0 / minsize - 1.0 / maxsize) / nact);
    for (ll i = nact - 1; i >= 0; i--) radial[nact - i] = h * maxsize / (h + maxsize * i * fs);
  }
  ll p2m(double x, double y) const { return point2morton(x, y, min_x, min_y, boxsize, level); }
  void m2box(ll m, ll& x, ll& y) const {
    if (m >= (1ll << (2 * level))) throw std::invalid_argument("Morton2Box");
    x = mergeFrom2((uint64_t)m);
    y = mergeFrom2((uint64_t)m >> 1);
  }
  bool isrel(ll m) const { return relboxes.count(m) > 0; }
  std::vector<ll> cousins(ll m) const {  // Level.h:261-289
    std::vector<ll> out;
    ll x, y;
    m2box(m, x, y);
    ll xlo = std::max<ll>((x / 2) * 2 - 2, 0), xup = std::min<ll>(nbps - 1, (x / 2) * 2 + 3);
    ll ylo = std::max<ll>((y / 2) * 2 - 2, 0), yup = std::min<ll>(nbps - 1, (y / 2) * 2 + 3);
    for (ll j = ylo; j <= yup; j++)
      for (ll i = xlo; i <= xup; i++)
        if (std::abs(i - x) > 1 || std::abs(j - y) > 1) {
          ll c = box2morton(i, j, level);
          if (isrel(c)) out.push_back(c);
        }
    std::sort(out.begin(), out.end());
    return out;
  }
  std::vector<ll> neighbours(ll m) const {  // Level.h:378-404
    std::vector<ll> out;
    ll x, y;
    m2box(m, x, y);
    ll xlo = std::max<ll>(0, x - 1), xup = std::min<ll>(nbps - 1, x + 1);
    ll ylo = std::max<ll>(0, y - 1), yup = std::min<ll>(nbps - 1, y + 1);
    for (ll j = ylo; j <= yup; j++)
      for (ll i = xlo; i <= xup; i++) {
        ll c = box2morton(i, j, level);
        if (isrel(c)) out.push_back(c);
      }
    std::sort(out.begin(), out.end());
    return out;
  }
  void center(ll m, double& x, double& y) const {  // :291-298
    ll px, py;
    m2box(m, px, py);
    x = min_x + px * boxsize + 0.5 * boxsize;
    y = min_y + py * boxsize + 0.5 * boxsize;
  }
  ll nrad() const { return (ll)radial.size() - 1; }
  ll cart2nonrel(ll m, double x, double y) const {  // :310-323
    double cxx, cyy;
    center(m, cxx, cyy);
    x -= cxx;
    y -= cyy;
    cart2pol(x, y);
    const ll posr = (ll)(std::upper_bound(radial.begin(), radial.end(), x) - radial.begin()) - 1;
    const ll post = (ll)(y * nce / M_PI);
    return post * nrad() + posr;
  }
  double chebposr2radius(ll posr, double s) const {  // :347-356
    double mn = radial[posr];
    double rmin = h / radial[posr + 1];
    double dr = h / mn - rmin;
    return h / ((s + 1.0) / 2.0 * dr + rmin);
  }
  double cheb2radius(ll nonrel, double s) const { return chebposr2radius(nonrel % nrad(), s); }
  void cheb2cart(ll m, ll nonrel, double& x, double& y) const {  // :358-375
    ll posr = nonrel % nrad(), post = nonrel / nrad();
    const double da = M_PI / nce;
    x = chebposr2radius(posr, x);
    y = (y + 1.0) / 2.0 * da + da * post;
    double cxx, cyy;
    center(m, cxx, cyy);
    pol2cart(x, y);
    x += cxx;
    y += cyy;
  }
  double polr2chebr(ll posr, double r) const {  // :413-422
    double mn = radial[posr];
    double rmin = h / radial[posr + 1];
    double dr = h / mn - rmin;
    return (h / r - rmin) / dr * 2.0 - 1.0;
  }
  void cart2cheb(ll m, ll nonrel, double& x, double& y) const {  // :424-437
    double cxx, cyy;
    center(m, cxx, cyy);
    x -= cxx;
    y -= cyy;
    cart2pol(x, y);
    ll posr = nonrel % nrad(), post = nonrel / nrad();
    const double da = M_PI / nce;
    x = polr2chebr(posr, x);
    y = (y - da * post) / da * 2.0 - 1.0;
  }
  void cart2cheb_find(ll m, double& x, double& y, ll& nonrel) const {  // :439-453
    double cxx, cyy;
    center(m, cxx, cyy);
    x -= cxx;
    y -= cyy;
    cart2pol(x, y);
    const ll posr = (ll)(std::upper_bound(radial.begin(), radial.end(), x) - radial.begin()) - 1;
    const ll post = (ll)(y * nce / M_PI);
    nonrel = post * nrad() + posr;
    const double da = M_PI / nce;
    x = polr2chebr(posr, x);
    y = (y - da * post) / da * 2.0 - 1.0;
  }
};

// Interpolator<Ps,Pang>, Interpolator.h:7-122, Chebyshev.h:10-68 (runtime sizes).
struct Interp {
  int Ps, Pang, P;
  std::vector<double> xs, xang;     // ascending Chebyshev points
  std::vector<double> Ts, Tang;     // Ts[m*Ps+k] = T_m(xs[k])
  Interp(int ps, int pa) : Ps(ps), Pang(pa), P(ps * pa) {
    auto pts = [](int n) {
      std::vector<double> x(n);
      for (int i = 0; i < n; i++) x[n - 1 - i] = std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * n));
      return x;
    };
    auto tn = [](const std::vector<double>& x) {
      int n = (int)x.size();
      std::vector<double> T(n * n);
      for (int i = 0; i < n; i++) {
        T[0 * n + i] = 1.0;
        if (n > 1) T[1 * n + i] = x[i];
      }
      for (int m = 2; m < n; m++)
        for (int i = 0; i < n; i++) T[m * n + i] = 2.0 * x[i] * T[(m - 1) * n + i] - T[(m - 2) * n + i];
      return T;
    };
    xs = pts(Ps);
    xang = pts(Pang);
    Ts = tn(xs);
    Tang = tn(xang);
  }
  void point(int it, double& x, double& y) const {
    x = xs[it % Ps];
    y = xang[it / Ps];
  }
  cd interpolate(double x, double y, const cd* v) const {  // Interpolator.h:55-77
    std::vector<double> tx(Ps), ty(Pang);
    tx[0] = 1.0;
    if (Ps > 1) tx[1] = x;
    for (int i = 2; i < Ps; i++) tx[i] = 2.0 * x * tx[i - 1] - tx[i - 2];
    ty[0] = 1.0;
    if (Pang > 1) ty[1] = y;
    for (int i = 2; i < Pang; i++) ty[i] = 2.0 * y * ty[i - 1] - ty[i - 2];
    cd res = 0.0;
    for (int a = 0; a < Pang; a++)
      for (int s = 0; s < Ps; s++) res += (ty[a] * tx[s]) * v[a * Ps + s];
    return res;
  }
  void generate(cd* arr) const {  // Interpolator.h:81-120
    std::vector<cd> C(P);
    for (int m = 0; m < Ps; m++)
      for (int n = 0; n < Pang; n++) {
        cd acc = 0.0;
        for (int k = 0; k < Ps; k++)
          for (int l = 0; l < Pang; l++) acc += arr[l * Ps + k] * Ts[m * Ps + k] * Tang[n * Pang + l];
        C[n * Ps + m] = acc;
      }
    for (int m = 0; m < Ps; m++)
      for (int n = 0; n < Pang; n++) {
        double an = n == 0 ? 1.0 : 2.0, am = m == 0 ? 1.0 : 2.0;
        arr[n * Ps + m] = C[n * Ps + m] * (am * an / (Ps * Pang));
      }
  }
};

struct Stats {
  ll near = 0, leaf = 0, prop = 0, interp = 0, cones = 0;
};

// BoxTree<Ps,Pang>, IFGF_Source/BoxTree.h:24-1171.
//
// Extension without a reference counterpart (the reference's targets are
// always its sources): an optional separate target set tx, ty. The bounding
// box then covers sources and targets, the targets are sorted by their leaf
// Morton code like the sources, and the three loops over targets (cones due
// to cousins, near field, interpolation) run over them instead of the
// sources; the near field skips no index (the kernel is zero for r < 1e-14).
// Without separate targets every step is the reference's.
struct BoxTree {
  std::vector<double> x_, y_, nx_, ny_;
  std::vector<ll> sorting_, inverse_;
  bool sep_ = false;                 // separate target set
  std::vector<double> tx_, ty_;      // sorted targets (the sorted sources if !sep_)
  std::vector<ll> tsorting_;
  ll M_ = 0;
  int nlevels_;
  double k_, eta_;
  int F_;
  ll N_;
  std::vector<std::unique_ptr<Level>> levels_;
  std::vector<cd> solution_, cur_, prev_;
  Interp ip_;
  Stats st_;

  BoxTree(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& nx,
          const std::vector<double>& ny, int nlev, double k, int F, double eta, int Ps, int Pang,
          const std::vector<double>* tx = nullptr, const std::vector<double>* ty = nullptr)
      : x_(x), y_(y), nx_(nx), ny_(ny), inverse_(x.size()), nlevels_(nlev), k_(k), eta_(eta),
        F_(F), N_((ll)x.size()), ip_(Ps, Pang) {
    if (nlevels_ < 3) throw std::invalid_argument("IFGF accelerator requires at least 3 levels.");
    if (tx && ty) {
      sep_ = true;
      tx_ = *tx;
      ty_ = *ty;
      M_ = (ll)tx_.size();
    }
    InitializeBoxesAndLevels();
    InitializeRelevantConeSegments();
    ll mx = 0;
    for (int l = 2; l < nlevels_; l++) mx = std::max<ll>(mx, (ll)levels_[l]->rel2nonrel.size());
    cur_.assign(mx * ip_.P, 0.0);
    prev_.assign(mx * ip_.P, 0.0);
    solution_.assign(M_, 0.0);
  }

  void InitializeBoxesAndLevels() {  // BoxTree.h:179-231
    double mn[2] = {1e308, 1e308}, mxv[2] = {-1e308, -1e308};
    for (ll i = 0; i < N_; i++) {
      mn[0] = std::min(mn[0], x_[i]);
      mn[1] = std::min(mn[1], y_[i]);
      mxv[0] = std::max(mxv[0], x_[i]);
      mxv[1] = std::max(mxv[1], y_[i]);
    }
    for (ll i = 0; i < M_; i++) {
      mn[0] = std::min(mn[0], tx_[i]);
      mn[1] = std::min(mn[1], ty_[i]);
      mxv[0] = std::max(mxv[0], tx_[i]);
      mxv[1] = std::max(mxv[1], ty_[i]);
    }
    mxv[0] += 1e-10;
    mxv[1] += 1e-10;
    mn[0] -= 1e-10;
    mn[1] -= 1e-10;
    const double bs = std::max(mxv[0] - mn[0], mxv[1] - mn[1]);
    // SortBox :73-106, stable sort by leaf Morton code (Vector.h:13-28)
    std::vector<ll> code(N_);
    sorting_.resize(N_);
    for (ll i = 0; i < N_; i++) {
      sorting_[i] = i;
      code[i] = point2morton(x_[i], y_[i], mn[0], mn[1], bs / (1ll << (nlevels_ - 1)), nlevels_ - 1);
    }
    std::stable_sort(sorting_.begin(), sorting_.end(),
                     [&](ll a, ll b) { return code[a] < code[b]; });
    auto arrange = [&](std::vector<double>& v) {
      std::vector<double> t(N_);
      for (ll i = 0; i < N_; i++) t[i] = v[sorting_[i]];
      v.swap(t);
    };
    arrange(x_);
    arrange(y_);
    arrange(nx_);
    arrange(ny_);
    for (ll i = 0; i < N_; i++) inverse_[sorting_[i]] = i;
    if (sep_) {
      std::vector<ll> tcode(M_);
      tsorting_.resize(M_);
      for (ll i = 0; i < M_; i++) {
        tsorting_[i] = i;
        tcode[i] = point2morton(tx_[i], ty_[i], mn[0], mn[1], bs / (1ll << (nlevels_ - 1)), nlevels_ - 1);
      }
      std::stable_sort(tsorting_.begin(), tsorting_.end(), [&](ll a, ll b) { return tcode[a] < tcode[b]; });
      std::vector<double> a(M_), b(M_);
      for (ll i = 0; i < M_; i++) {
        a[i] = tx_[tsorting_[i]];
        b[i] = ty_[tsorting_[i]];
      }
      tx_.swap(a);
      ty_.swap(b);
    } else {
      tx_ = x_;
      ty_ = y_;
      tsorting_ = sorting_;
      M_ = N_;
    }
    levels_.resize(nlevels_);
    for (int i = 1; i < nlevels_; i++)
      levels_[i].reset(new Level(i, nlevels_, mn[0], mn[1], bs / (1ll << i), k_));
    // InitializeLevelDBoxesData :111-146
    Level& L = *levels_.back();
    ll old = -1, np = -1;
    for (ll i = 0; i < N_; i++) {
      ll m = L.p2m(x_[i], y_[i]);
      if (m != old) {
        if (i != 0) L.box2pts[old] = {i - np, np};
        old = m;
        np = 1;
        L.relboxes.insert(m);
      } else {
        np++;
      }
    }
    L.box2pts[old] = {N_ - np, np};
    // InitializeRelBoxesAllLevels :149-174
    std::unordered_set<ll> rel = L.relboxes, par;
    for (int i = nlevels_ - 1; i >= 2; i--) {
      if (i != nlevels_ - 1) levels_[i]->relboxes.insert(rel.begin(), rel.end());
      for (ll m : rel) par.insert(m / 4);
      std::swap(rel, par);
      par.clear();
    }
  }

  // :235-277
  void ConesDueToCousins(int level, std::map<ll, std::vector<ll>>& out) const {
    const Level& L = *levels_[level];
    ll old = -1;
    std::vector<ll> cz;
    std::map<ll, std::set<ll>> tmp;
    for (ll i = 0; i < M_; i++) {
      ll m = L.p2m(tx_[i], ty_[i]);
      if (m != old) {
        cz = L.cousins(m);
        old = m;
      }
      for (ll c : cz) tmp[c].insert(L.cart2nonrel(c, tx_[i], ty_[i]));
    }
    for (auto& kv : tmp) out[kv.first] = std::vector<ll>(kv.second.begin(), kv.second.end());
  }

  std::vector<ll> RelChildren(int level, ll m) const {  // :279-297
    std::vector<ll> ch;
    if (level >= nlevels_ - 1) return ch;
    for (ll c = m * 4; c < (m + 1) * 4; c++)
      if (levels_[level + 1]->isrel(c)) ch.push_back(c);
    return ch;
  }

  // :299-428
  void ConesDueToParents(int level, std::map<ll, std::vector<ll>>& out) {
    Level& P = *levels_[level - 1];
    Level& L = *levels_[level];
    std::map<ll, std::set<ll>> all;
    ll oldb = -1;
    std::vector<ll> ch;
    for (size_t it = 0; it < P.rel2nonrel.size(); it++) {
      const ll pb = P.rel2box[it], q = P.rel2nonrel[it];
      if (pb != oldb) {
        ch = RelChildren(level - 1, pb);
        if (ch.empty()) throw std::logic_error("Every cocentered morton box must have children");
        oldb = pb;
      }
      for (ll c : ch) {
        const ll cp = c % 4;
        auto f = L.protobox.find(q);
        if (f == L.protobox.end() || f->second[cp].empty()) {
          auto& slot = L.protobox[q][cp];
          for (int j = 0; j < ip_.P; j++) {
            double x, y;
            ip_.point(j, x, y);
            P.cheb2cart(pb, q, x, y);
            slot[L.cart2nonrel(c, x, y)].push_back(j);
          }
        }
      }
    }
    oldb = -1;
    for (size_t it = 0; it < P.rel2nonrel.size(); it++) {
      const ll pb = P.rel2box[it], q = P.rel2nonrel[it];
      if (pb != oldb) {
        ch = RelChildren(level - 1, pb);
        oldb = pb;
      }
      for (ll c : ch)
        for (auto& kv : L.protobox.at(q)[c % 4]) all[c].insert(kv.first);
    }
    for (auto& kv : all) out[kv.first] = std::vector<ll>(kv.second.begin(), kv.second.end());
  }

  void InitializeRelevantConeSegments() {  // :553-583, merge :430-459, flatten :461-514
    for (int level = 2; level < nlevels_; level++) {
      std::map<ll, std::vector<ll>> ic, pc;
      ConesDueToCousins(level, ic);
      if (level > 2) ConesDueToParents(level, pc);
      if (ic.empty() && pc.empty() && !sep_)  // separate targets: coarse levels may have none
        throw std::logic_error("The number of cone segments on any level cannot be 0");
      std::map<ll, std::set<ll>> merged;
      for (auto& kv : ic) merged[kv.first].insert(kv.second.begin(), kv.second.end());
      for (auto& kv : pc) merged[kv.first].insert(kv.second.begin(), kv.second.end());
      Level& L = *levels_[level];
      for (auto& kv : merged)
        for (ll q : kv.second) {
          L.boxnonrel2rel[kv.first][q] = (ll)L.rel2nonrel.size();
          L.rel2nonrel.push_back(q);
          L.rel2box.push_back(kv.first);
        }
      st_.cones += (ll)L.rel2nonrel.size();
    }
  }

  // SingularInteractions<F>, :614-683
  // Targets are independent: with OpenMP each thread takes a contiguous
  // range of targets and sums every target in the serial order, so the
  // result does not depend on the thread count.
  void Near(const std::vector<cd>& dens, const std::vector<std::unordered_set<ll>>* excl) {
    Level& L = *levels_.back();
    ll cnt = 0;
    bool bad = false;
#pragma omp parallel reduction(+ : cnt) reduction(|| : bad)
    {
      ll old = -1;
      std::vector<ll> nb;
#pragma omp for schedule(static)
      for (ll i = 0; i < M_; i++) {
        ll m = L.p2m(tx_[i], ty_[i]);
        cd res = solution_[i];
        if (m != old) {
          nb = L.neighbours(m);
          if (nb.size() > 9) bad = true;
          old = m;
        }
        for (ll b : nb) {
          const auto& pd = L.box2pts.at(b);
          for (ll s = pd[0]; s < pd[0] + pd[1]; s++) {
            if (!sep_ && s == i) continue;
            if (excl && (*excl)[i].count(s)) continue;
            res += kernel(F_, tx_[i], ty_[i], x_[s], y_[s], nx_[s], ny_[s], k_, eta_) * dens[s];
            cnt++;
          }
        }
        solution_[i] = res;
      }
    }
    if (bad) throw std::logic_error("Cannot have more than 9 neighbours.");
    st_.near += cnt;
  }

  // LevelDEvaluations<F>, :685-762
  void Leaf(const std::vector<cd>& dens) {
    Level& L = *levels_.back();
    ll cnt = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : cnt)
    for (size_t c = 0; c < L.rel2nonrel.size(); c++) {
      const ll b = L.rel2box[c], q = L.rel2nonrel[c];
      const auto& pd = L.box2pts.at(b);
      for (int j = 0; j < ip_.P; j++) {
        double x, y;
        ip_.point(j, x, y);
        const double radius = L.cheb2radius(q, x);
        L.cheb2cart(b, q, x, y);
        cd res = 0.0;
        for (ll s = pd[0]; s < pd[0] + pd[1]; s++) {
          res += kernel(F_, x, y, x_[s], y_[s], nx_[s], ny_[s], k_, eta_) * dens[s];
          cnt++;
        }
        prev_[c * ip_.P + j] = res / factorization(radius, k_);
      }
      ip_.generate(&prev_[c * ip_.P]);
    }
    st_.leaf += cnt;
    cur_ = prev_;
  }

  // Propagation(level), :767-868
  void Propagation(int level) {
    Level& P = *levels_[level - 1];
    Level& L = *levels_[level];
    const int NP = ip_.P;
    ll cnt = 0;
#pragma omp parallel reduction(+ : cnt)
    {
    std::vector<double> xs(NP), ys(NP), radii(NP);
    std::vector<cd> val(NP);
    ll oldb = -1;
    std::vector<ll> ch;
#pragma omp for schedule(static)
    for (size_t it = 0; it < P.rel2nonrel.size(); it++) {
      const ll pb = P.rel2box[it], q = P.rel2nonrel[it];
      if (pb != oldb) {
        ch = RelChildren(level - 1, pb);
        oldb = pb;
      }
      const auto& proto = L.protobox.at(q);
      for (int j = 0; j < NP; j++) {
        ip_.point(j, xs[j], ys[j]);
        radii[j] = P.cheb2radius(q, xs[j]);
        P.cheb2cart(pb, q, xs[j], ys[j]);
        val[j] = 0.0;
      }
      for (ll c : ch) {
        for (auto& kv : proto[c % 4]) {
          const ll cq = kv.first;
          const ll crel = L.boxnonrel2rel.at(c).at(cq);
          const cd* coeffs = &prev_[crel * NP];
          for (int j : kv.second) {
            double lx = xs[j], ly = ys[j];
            L.cart2cheb(c, cq, lx, ly);
            const double lr = L.cheb2radius(cq, lx);
            val[j] += ip_.interpolate(lx, ly, coeffs) * factorization(lr, k_);
            cnt++;
          }
        }
      }
      for (int j = 0; j < NP; j++) cur_[it * NP + j] = val[j] / factorization(radii[j], k_);
      ip_.generate(&cur_[it * NP]);
    }
    }
    st_.prop += cnt;
  }

  // Interpolation(level), :876-937
  void Interpolation(int level) {
    Level& L = *levels_[level];
    ll cnt = 0;
#pragma omp parallel reduction(+ : cnt)
    {
    ll old = -1;
    std::vector<ll> cz;
#pragma omp for schedule(static)
    for (ll i = 0; i < M_; i++) {
      const double x = tx_[i], y = ty_[i];
      cd v = solution_[i];
      ll m = L.p2m(x, y);
      if (m != old) {
        cz = L.cousins(m);
        old = m;
      }
      for (ll c : cz) {
        double lx = x, ly = y;
        ll q;
        L.cart2cheb_find(c, lx, ly, q);
        const ll rel = L.boxnonrel2rel.at(c).at(q);
        const double radius = L.cheb2radius(q, lx);
        v += ip_.interpolate(lx, ly, &prev_[rel * ip_.P]) * factorization(radius, k_);
        cnt++;
      }
      solution_[i] = v;
    }
    }
    st_.interp += cnt;
  }

  // Solve<F>, :1020-1071. dens in original order, overwritten with the result
  // (with separate targets: resized to the targets, in their input order).
  void Solve(bool with_near, std::vector<cd>& dens, const std::vector<std::unordered_set<ll>>* excl) {
    st_.near = st_.leaf = st_.prop = st_.interp = 0;
    std::fill(solution_.begin(), solution_.end(), cd(0.0));
    std::vector<cd> d(N_);
    for (ll i = 0; i < N_; i++) d[i] = dens[sorting_[i]];
    if (with_near) Near(d, excl);
    Leaf(d);
    for (int level = nlevels_ - 1; level >= 2; level--) {
      if (level > 2) Propagation(level);
      Interpolation(level);
      std::swap(cur_, prev_);
    }
    dens.resize(M_);
    for (ll i = 0; i < M_; i++) dens[tsorting_[i]] = solution_[i];
  }
};

// ----------------------------------------------------------------------------
// ForwardMap<Np, F, Ps, Pang>, ForwardMap/ForwardMap.hpp:49-177, .tpp.
struct Patch {
  double t1, t2;
  int curve;
  std::vector<double> tvals, jac;
  std::vector<ll> ns_idx;
  std::vector<double> ns_t;
  std::unordered_set<ll> ns_lookup;
  std::vector<ll> ns_est;
  double bb[4];                // xmin, xmax, ymin, ymax (+- delta)
  std::vector<cd> precomp;     // Np x (Np + n_ns), column-major like Eigen
};

struct ForwardMap {
  int Np, F;
  double k, eta, delta, p;
  int Ps, Pang, nlev;
  double ident;  // 1/2 for the exterior double-layer / combined-field operator
  std::vector<Curve> curves;
  std::vector<Patch> patches;
  std::vector<double> xs, ys, nxs, nys;
  std::vector<double> fn, fw, fnns, fwns;
  std::vector<std::unordered_set<ll>> sort_sing;
  std::unique_ptr<BoxTree> boxes;
  Stats last;
  double cheb[64][64];  // T_i(x_k) at the Np Fejer nodes, Chebyshev1d.hpp:134-163

  ForwardMap(const std::vector<Curve>& cv, const std::vector<std::vector<double>>& lims, int np,
             int f, double kk, double et, double dl, double pp, int ps, int pa, int nl)
      : Np(np), F(f), k(kk), eta(et), delta(dl), p(pp), Ps(ps), Pang(pa), nlev(nl), curves(cv) {
    ident = (F == kSingleLayer) ? 0.0 : 0.5;
    fejer1(Np, fn, fw);        // ForwardMap.tpp:22-28
    fejer1(40, fnns, fwns);    // num_ns = 40, ForwardMap.hpp:70
    if (Np > 64) throw std::invalid_argument("Np > 64");
    // recurrence form as the reference (setup_multiple_Tn, Chebyshev1d.hpp:84-110)
    for (int kk2 = 0; kk2 < Np; kk2++) {
      const double x = std::cos(M_PI * (2.0 * kk2 + 1.0) / (2.0 * Np));
      cheb[0][kk2] = 1.0;
      if (Np > 1) cheb[1][kk2] = x;
      for (int i = 2; i < Np; i++) cheb[i][kk2] = 2.0 * x * cheb[i - 1][kk2] - cheb[i - 2][kk2];
    }
    init_points_and_patches(lims);
    determine_near_singular();
  }

  // init_points_and_patches, ForwardMap.tpp:37-222 (closed curves).
  void init_points_and_patches(const std::vector<std::vector<double>>& lims) {
    int start = 0;
    for (size_t c = 0; c < curves.size(); c++) {
      const int np = (int)lims[c].size() - 1;
      for (int pi = 0; pi < np; pi++) {
        Patch P;
        P.t1 = lims[c][pi];
        P.t2 = lims[c][pi + 1];
        P.curve = (int)c;
        for (int j = 0; j < Np; j++) {
          const double t = ab2cd(fn[j], -1.0, 1.0, P.t1, P.t2);
          P.tvals.push_back(t);
          const double n0 = curves[c].ypt(t), n1 = -curves[c].xpt(t);
          const double nn = std::sqrt(n0 * n0 + n1 * n1);
          P.jac.push_back(nn);
          xs.push_back(curves[c].xt(t));
          ys.push_back(curves[c].yt(t));
          nxs.push_back(n0 / nn);
          nys.push_back(n1 / nn);
        }
        comp_bounding_box(P);
        // adjacent patches, closed curve (:164-218)
        if (pi == 0) {
          P.ns_est = {start + np - 1, start + pi + 1};
        } else if (pi == np - 1) {
          P.ns_est = {start + pi - 1, start};
        } else {
          P.ns_est = {start + pi - 1, start + pi + 1};
        }
        patches.push_back(std::move(P));
      }
      start += np;
    }
  }

  // Patch::comp_bounding_box, Patch.tpp:36-64 (extrema located by sign changes
  // of x'(t), y'(t) on a fine sample, refined by bisection).
  void comp_bounding_box(Patch& P) {
    const Curve& cv = curves[P.curve];
    double xmn = std::min(cv.xt(P.t1), cv.xt(P.t2)), xmx = std::max(cv.xt(P.t1), cv.xt(P.t2));
    double ymn = std::min(cv.yt(P.t1), cv.yt(P.t2)), ymx = std::max(cv.yt(P.t1), cv.yt(P.t2));
    const int ns = 64;
    for (int which = 0; which < 2; which++) {
      auto d = [&](double t) { return which == 0 ? cv.xpt(t) : cv.ypt(t); };
      for (int i = 0; i < ns; i++) {
        double a = P.t1 + (P.t2 - P.t1) * i / ns, b = P.t1 + (P.t2 - P.t1) * (i + 1) / ns;
        double fa = d(a), fb = d(b);
        if (fa == 0.0 || fa * fb < 0.0) {
          for (int it = 0; it < 80 && fa != 0.0; it++) {
            double m = 0.5 * (a + b), fm = d(m);
            if ((fa < 0) == (fm < 0)) {
              a = m;
              fa = fm;
            } else {
              b = m;
            }
          }
          const double tr = a;
          if (which == 0) {
            xmn = std::min(xmn, cv.xt(tr));
            xmx = std::max(xmx, cv.xt(tr));
          } else {
            ymn = std::min(ymn, cv.yt(tr));
            ymx = std::max(ymx, cv.yt(tr));
          }
        }
      }
    }
    P.bb[0] = xmn - delta;
    P.bb[1] = xmx + delta;
    P.bb[2] = ymn - delta;
    P.bb[3] = ymx + delta;
  }

  // determine_patch_near_singular_points, ForwardMap.tpp:224-285.
  void determine_near_singular() {
    for (auto& P : patches) {
      const Curve& cv = curves[P.curve];
      for (ll tp : P.ns_est) {
        for (ll kk2 = tp * Np; kk2 < tp * Np + Np; kk2++) {
          const double xp = xs[kk2], yp = ys[kk2];
          if (!(xp >= P.bb[0] && xp <= P.bb[1] && yp >= P.bb[2] && yp <= P.bb[3])) continue;
          auto f = [&](double t) {
            const double dx = cv.xt(t) - xp, dy = cv.yt(t) - yp;
            return std::sqrt(std::abs(dx * dx + dy * dy));
          };
          double ts;
          if (f(P.t1) < delta) {
            ts = P.t1;
          } else if (f(P.t2) < delta) {
            ts = P.t2;
          } else {
            double a = P.t1, b = P.t2;
            const double g = 0.5 * (std::sqrt(5.0) - 1.0);
            double c1 = b - g * (b - a), c2 = a + g * (b - a), f1 = f(c1), f2 = f(c2);
            for (int it = 0; it < 200 && (b - a) > 1e-12 * (1.0 + std::abs(a)); it++) {
              if (f1 < f2) {
                b = c2;
                c2 = c1;
                f2 = f1;
                c1 = b - g * (b - a);
                f1 = f(c1);
              } else {
                a = c1;
                c1 = c2;
                f1 = f2;
                c2 = a + g * (b - a);
                f2 = f(c2);
              }
            }
            ts = 0.5 * (a + b);
            if (!(f(ts) < delta)) continue;
          }
          P.ns_idx.push_back(kk2);
          P.ns_lookup.insert(kk2);
          P.ns_t.push_back(ts);
        }
      }
    }
  }

  // single_patch_point_compute_precomputations, ForwardMap.tpp:287-389. The
  // target is the curve point at parameter ttgt (its own node, or the
  // near-singular node of the adjacent patch).
  void precomp_column(const Patch& P, double tsing, double ttgt, cd* out) const {
    const Curve& cv = curves[P.curve];
    const int n = 40;
    const double alpha = ab2cd(tsing, P.t1, P.t2, -1.0, 1.0);
    int loc = 0;
    if (std::abs(-1.0 - alpha) < 1e-5) loc = -1;
    else if (std::abs(1.0 - alpha) < 1e-5) loc = 1;
    const double ts = loc == 0 ? tsing : (loc < 0 ? P.t1 : P.t2);
    double tt = ttgt - ts;
    if (tt > M_PI) tt -= 2.0 * M_PI;
    if (tt < -M_PI) tt += 2.0 * M_PI;
    std::vector<double> xi(n);
    std::vector<cd> integ(n);
    for (int i = 0; i < n; i++) {
      double dxi, off;
      if (loc == 0) {
        xi[i] = xi_cov_center(fnns[i], alpha, p);
        off = ((sgn(fnns[i]) - alpha) / M_PI) * w_cov_base(M_PI * std::abs(fnns[i]), p);
        dxi = dxi_cov_center(fnns[i], alpha, p);
      } else if (loc == -1) {
        xi[i] = xi_cov_left(fnns[i], p);
        off = (2.0 / M_PI) * w_cov_base(M_PI * std::abs((fnns[i] + 1.0) / 2.0), p);
        dxi = dxi_cov_left(fnns[i], p);
      } else {
        xi[i] = xi_cov_right(fnns[i], p);
        off = -(2.0 / M_PI) * w_cov_base(M_PI * std::abs((fnns[i] - 1.0) / 2.0), p);
        dxi = dxi_cov_right(fnns[i], p);
      }
      const double g = ab2cd(xi[i], -1.0, 1.0, P.t1, P.t2);
      const double n0 = cv.ypt(g), n1 = -cv.xpt(g);
      const double cj = std::sqrt(n0 * n0 + n1 * n1);
      double dx, dy;
      cv.diff(ttgt, tt - ((P.t2 - P.t1) / 2.0) * off, dx, dy);
      const cd gr = kernel_d(F, dx, dy, n0 / cj, n1 / cj, k, eta);
      integ[i] = ((P.t2 - P.t1) / 2.0) * gr * cj * dxi * fwns[i];
    }
    std::vector<double> tm1(n, 1.0), tc(xi), tp1(n);
    cd s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < n; i++) {
      s0 += integ[i];
      s1 += integ[i] * tc[i];
    }
    out[0] = s0;
    if (Np > 1) out[1] = s1;
    for (int d = 2; d < Np; d++) {
      cd s = 0.0;
      for (int i = 0; i < n; i++) {
        tp1[i] = 2.0 * xi[i] * tc[i] - tm1[i];
        s += integ[i] * tp1[i];
      }
      out[d] = s;
      std::swap(tm1, tc);
      std::swap(tc, tp1);
    }
  }

  double tval_of(ll i) const { return patches[i / Np].tvals[i % Np]; }

  // compute_precomputations, ForwardMap.tpp:391-423.
  void compute_precomputations() {
#pragma omp parallel for schedule(dynamic, 8)
    for (size_t pi = 0; pi < patches.size(); pi++) {
      Patch& P = patches[pi];
      const size_t ncol = Np + P.ns_t.size();
      P.precomp.assign(Np * ncol, 0.0);
      for (int j = 0; j < Np; j++)
        precomp_column(P, P.tvals[j], P.tvals[j], &P.precomp[j * Np]);
      for (size_t j = 0; j < P.ns_idx.size(); j++)
        precomp_column(P, P.ns_t[j], tval_of(P.ns_idx[j]), &P.precomp[(Np + j) * Np]);
    }
  }

  // precomps_and_setup, ForwardMap.tpp:542-565 (+ CheckIfSingandNearSingInNeighborhood,
  // BoxTree.h:1079-1161, and init_sort_sing_point, ForwardMap.tpp:567-607).
  void setup() {
    compute_precomputations();
    boxes.reset(new BoxTree(xs, ys, nxs, nys, nlev, k, F, eta, Ps, Pang));
    check_neighbourhood();
    const auto& inv = boxes->inverse_;
    const ll N = (ll)xs.size();
    sort_sing.assign(N, {});
    for (size_t pi = 0; pi < patches.size(); pi++) {
      std::unordered_set<ll> s;
      for (ll i = pi * Np; i < (ll)(pi + 1) * Np; i++) s.insert(inv[i]);
      for (ll i = pi * Np; i < (ll)(pi + 1) * Np; i++) sort_sing[inv[i]] = s;
    }
    for (size_t pi = 0; pi < patches.size(); pi++)
      for (ll i : patches[pi].ns_idx)
        for (ll j = pi * Np; j < (ll)(pi + 1) * Np; j++) sort_sing[inv[i]].insert(inv[j]);
  }

  void check_neighbourhood() {
    const Level& L = *boxes->levels_.back();
    auto in_nb = [&](ll i, ll j) {
      const auto nb = L.neighbours(L.p2m(xs[i], ys[i]));
      const ll m = L.p2m(xs[j], ys[j]);
      return std::find(nb.begin(), nb.end(), m) != nb.end();
    };
    for (ll i = 0; i < (ll)xs.size(); i++) {
      const ll ps = (i / Np) * Np;
      for (ll j = ps; j < ps + Np; j++)
        if (j != i && !in_nb(i, j)) throw std::logic_error("Singular point not in neighborhood");
    }
    for (size_t pi = 0; pi < patches.size(); pi++)
      for (ll i : patches[pi].ns_idx)
        for (ll j = pi * Np; j < (ll)(pi + 1) * Np; j++)
          if (!in_nb(i, j)) throw std::logic_error("Near singular point not in neighborhood");
  }

  // compute_sing_near_sing_interactions, ForwardMap.tpp:481-515 with
  // Cheb1D::interp_1d<Np>, Chebyshev1d.hpp:134-163.
  std::vector<cd> sing_near_sing(const std::vector<cd>& dens) const {
    std::vector<cd> out(dens.size(), 0.0);
    std::vector<cd> c(Np);
    for (size_t pi = 0; pi < patches.size(); pi++) {
      const Patch& P = patches[pi];
      const cd* d = &dens[pi * Np];
      for (int i = 0; i < Np; i++) {
        cd s = 0.0;
        for (int kk2 = 0; kk2 < Np; kk2++) s += d[kk2] * cheb[i][kk2];
        c[i] = s * (i == 0 ? 1.0 / Np : 2.0 / Np);
      }
      const size_t ncol = Np + P.ns_idx.size();
      for (size_t col = 0; col < ncol; col++) {
        cd s = 0.0;
        for (int i = 0; i < Np; i++) s += c[i] * P.precomp[col * Np + i];
        if ((int)col < Np) out[pi * Np + col] += s;
        else out[P.ns_idx[col - Np]] += s;
      }
    }
    return out;
  }

  // compute_intensities, ForwardMap.tpp:425-435.
  void intensities(std::vector<cd>& d) const {
    for (size_t pi = 0; pi < patches.size(); pi++) {
      const double j = ab2cdjac(-1, 1, patches[pi].t1, patches[pi].t2);
      for (int q = 0; q < Np; q++) d[pi * Np + q] *= j * fw[q] * patches[pi].jac[q];
    }
  }

  // compute_Ax_acc, ForwardMap.tpp:609-655 (+ the 1/2 identity of the
  // exterior double-layer jump for DoubleLayer / Combined).
  std::vector<cd> apply(const std::vector<cd>& x) {
    std::vector<cd> sns = sing_near_sing(x);
    std::vector<cd> d(x);
    intensities(d);
    boxes->Solve(true, d, &sort_sing);
    last = boxes->st_;
    for (size_t i = 0; i < x.size(); i++) sns[i] += d[i] + ident * x[i];
    return sns;
  }

  // compute_Ax_unacc, ForwardMap.tpp:438-540 (O(N^2) direct sum + correction).
  std::vector<cd> apply_unacc(const std::vector<cd>& x) const {
    std::vector<cd> sns = sing_near_sing(x);
    std::vector<cd> d(x);
    intensities(d);
    const ll np = (ll)patches.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (ll p1 = 0; p1 < np; p1++)
      for (ll xi = p1 * Np; xi < (p1 + 1) * Np; xi++) {
        cd acc = 0.0;
        for (ll p2 = 0; p2 < np; p2++) {
          if (p1 == p2 || patches[p2].ns_lookup.count(xi)) continue;
          for (ll s = p2 * Np; s < (p2 + 1) * Np; s++)
            acc += kernel(F, xs[xi], ys[xi], xs[s], ys[s], nxs[s], nys[s], k, eta) * d[s];
        }
        sns[xi] += acc + ident * x[xi];
      }
    return sns;
  }

  // propagate_unacc, ForwardMap.tpp:702-727.
  std::vector<cd> propagate(const std::vector<double>& tx, const std::vector<double>& ty,
                            const std::vector<cd>& dens) const {
    std::vector<cd> d(dens);
    intensities(d);
    std::vector<cd> out(tx.size(), 0.0);
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t i = 0; i < tx.size(); i++)
      for (size_t j = 0; j < xs.size(); j++)
        out[i] += d[j] * kernel(F, tx[i], ty[i], xs[j], ys[j], nxs[j], nys[j], k, eta);
    return out;
  }
};

// ----------------------------------------------------------------------------
// Householder GMRES, ForwardMap/opGMRES.hpp:68-269 (Walker 1988), identity
// preconditioner; Householder reflectors and Givens rotations in the
// convention of Eigen's makeHouseholder / JacobiRotation::makeGivens.
// makeHouseholder builds a unitary, non-Hermitian H with H r = beta e1 for
// complex r, so the Krylov vectors (opGMRES.hpp:147-149) and the solution
// update (:224-229) are formed with H^H = H^-1; applying H there as written
// gives a consistent but not residual-minimising iteration for complex data.
struct House {
  std::vector<cd> ess;  // essential part (length m - i - 1)
  cd tau = 0.0;
};
static double cabs2(cd z) { return std::norm(z); }
// makeHouseholder on v (length n): v -> beta e1, H = I - tau [1;ess][1;ess]^H
static void make_householder(const cd* v, int n, House& h, double& beta) {
  double tail = 0.0;
  for (int i = 1; i < n; i++) tail += cabs2(v[i]);
  const cd c0 = v[0];
  h.ess.assign(n > 0 ? n - 1 : 0, 0.0);
  const double tol = std::numeric_limits<double>::min();
  if (tail <= tol && c0.imag() * c0.imag() <= tol) {
    h.tau = 0.0;
    beta = c0.real();
  } else {
    beta = std::sqrt(cabs2(c0) + tail);
    if (c0.real() >= 0.0) beta = -beta;
    for (int i = 1; i < n; i++) h.ess[i - 1] = v[i] / (c0 - beta);
    h.tau = std::conj((beta - c0) / beta);
  }
}
// applyHouseholderOnTheLeft on v (length n): v <- H v, H = I - tau [1;ess][1;ess]^H
// (adjoint = true applies H^H = H^-1, i.e. conj(tau)).
static void apply_householder(cd* v, int n, const House& h, bool adjoint = false) {
  const cd tau = adjoint ? std::conj(h.tau) : h.tau;
  if (n == 1) {
    v[0] *= (1.0 - tau);
    return;
  }
  if (tau == 0.0) return;
  cd t = v[0];
  for (int i = 1; i < n; i++) t += std::conj(h.ess[i - 1]) * v[i];
  v[0] -= tau * t;
  for (int i = 1; i < n; i++) v[i] -= tau * h.ess[i - 1] * t;
}
struct Givens {
  cd c = 1.0, s = 0.0;
};
static Givens make_givens(cd p, cd q) {
  Givens g;
  if (q == 0.0) {
    g.c = p.real() < 0 ? -1.0 : 1.0;
    g.s = 0.0;
  } else if (p == 0.0) {
    g.c = 0.0;
    g.s = -q / std::abs(q);
  } else {
    const double p1 = std::abs(p.real()) + std::abs(p.imag());
    const double q1 = std::abs(q.real()) + std::abs(q.imag());
    if (p1 >= q1) {
      cd ps = p / p1, qs = q / p1;
      double p2 = cabs2(ps), q2 = cabs2(qs);
      double u = std::sqrt(1.0 + q2 / p2);
      if (p.real() < 0) u = -u;
      g.c = 1.0 / u;
      g.s = -qs * std::conj(ps) * (g.c.real() / p2);
    } else {
      cd ps = p / q1, qs = q / q1;
      double p2 = cabs2(ps), q2 = cabs2(qs);
      double u = q1 * std::sqrt(p2 + q2);
      if (p.real() < 0) u = -u;
      double pa = std::abs(p);
      ps = p / pa;
      g.c = pa / u;
      g.s = -std::conj(ps) * (q / u);
    }
  }
  return g;
}
// rows (a, b) <- G^* applied on the left
static void apply_givens_adj(cd& a, cd& b, const Givens& g) {
  const cd c = std::conj(g.c), s = -g.s;  // adjoint
  const cd x = a, y = b;
  a = c * x + std::conj(s) * y;
  b = -s * x + std::conj(c) * y;
}

// Returns true on convergence; iters / err updated like opGMRES.hpp:382-392.
bool gmres(const std::function<void(const std::vector<cd>&, std::vector<cd>&)>& A,
           const std::vector<cd>& rhs, std::vector<cd>& x, int restart, double tol, int maxit,
           int& iters, double& err, std::vector<double>* hist) {
  const int m = (int)rhs.size();
  double rn = 0.0;
  for (auto& v : rhs) rn += std::norm(v);
  if (std::sqrt(rn) <= std::numeric_limits<double>::min()) {
    std::fill(x.begin(), x.end(), cd(0.0));
    err = 0.0;
    iters = 0;
    return true;
  }
  iters = 0;
  std::vector<cd> t(m), v(m), r0(m), xn(m);
  A(x, t);
  for (int i = 0; i < m; i++) r0[i] = rhs[i] - t[i];
  double r0n = 0.0;
  for (auto& q : r0) r0n += std::norm(q);
  r0n = std::sqrt(r0n);
  if (r0n == 0.0) {
    err = 0.0;
    return true;
  }
  std::vector<House> H(restart + 1);
  std::vector<std::vector<cd>> R(restart + 1, std::vector<cd>(restart + 1, 0.0));
  std::vector<cd> w(restart + 1, 0.0);
  std::vector<Givens> G(restart);
  double beta;
  make_householder(r0.data(), m, H[0], beta);
  w[0] = beta;
  for (int k = 1; k <= restart; ++k) {
    ++iters;
    std::fill(v.begin(), v.end(), cd(0.0));
    v[k - 1] = 1.0;
    for (int i = k - 1; i >= 0; --i) apply_householder(&v[i], m - i, H[i], true);
    A(v, t);
    v = t;
    for (int i = 0; i < k; ++i) apply_householder(&v[i], m - i, H[i]);
    double tn = 0.0;
    for (int i = k; i < m; i++) tn += std::norm(v[i]);
    if (tn != 0.0 && k < m) {
      make_householder(&v[k], m - k, H[k], beta);
      apply_householder(&v[k], m - k, H[k]);
    }
    for (int i = 0; i < k - 1; ++i) apply_givens_adj(v[i], v[i + 1], G[i]);
    if (k < m && v[k] != 0.0) {
      G[k - 1] = make_givens(v[k - 1], v[k]);
      apply_givens_adj(v[k - 1], v[k], G[k - 1]);
      apply_givens_adj(w[k - 1], w[k], G[k - 1]);
    }
    for (int i = 0; i < k; i++) R[i][k - 1] = v[i];
    err = std::abs(w[k]) / r0n;
    if (hist) hist->push_back(err);
    const bool stop = (k == m || err < tol || iters == maxit);
    if (stop || k == restart) {
      std::vector<cd> y(w.begin(), w.begin() + k);
      for (int i = k - 1; i >= 0; --i) {
        for (int j = i + 1; j < k; j++) y[i] -= R[i][j] * y[j];
        y[i] /= R[i][i];
      }
      std::fill(xn.begin(), xn.end(), cd(0.0));
      for (int i = k - 1; i >= 0; --i) {
        xn[i] += y[i];
        apply_householder(&xn[i], m - i, H[i], true);
      }
      for (int i = 0; i < m; i++) x[i] += xn[i];
      if (stop) return true;
      k = 0;
      A(x, t);
      for (int i = 0; i < m; i++) r0[i] = rhs[i] - t[i];
      for (auto& h : H) {
        h.ess.clear();
        h.tau = 0.0;
      }
      for (auto& row : R) std::fill(row.begin(), row.end(), cd(0.0));
      std::fill(w.begin(), w.end(), cd(0.0));
      make_householder(r0.data(), m, H[0], beta);
      w[0] = beta;
    }
  }
  return false;
}

}  // namespace orc

// ============================================================================
// C ABI for the test harness (ctypes). Status 0 = ok, nonzero = error; the
// message is available from oracle_last_error().
// ============================================================================
using namespace orc;

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// Threads of the oracle's OpenMP loops (n < 1: the runtime default); returns
// the count now in effect. Every loop sums each output in the serial order,
// so the results are the same for any count.
int oracle_set_threads(int n) {
  if (n >= 1) omp_set_num_threads(n);
  return omp_get_max_threads();
}

// The benchmark's counter-based generator: splitmix64 value i of a seed
// (value i = mix(seed + (start + i + 1) * 0x9E3779B97F4A7C15)), and the
// seeded complex density built from it by Box-Muller (entry i of column c
// from values 2m, 2m + 1, m = c n + i; u = ((v >> 11) + 1) / 2^53).
static unsigned long long splitmix_at(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void oracle_splitmix(unsigned long long seed, long long n, long long start, unsigned long long* out) {
  for (long long i = 0; i < n; i++) out[i] = splitmix_at(seed, (unsigned long long)(start + i));
}

void oracle_seeded_density(unsigned long long seed, long long n, int column, double* out) {
  const unsigned long long base = 2ull * (unsigned long long)column * (unsigned long long)n;
  for (long long i = 0; i < n; i++) {
    const double u1 = (double)((splitmix_at(seed, base + 2 * i) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (double)((splitmix_at(seed, base + 2 * i + 1) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double r = std::sqrt(-2.0 * std::log(u1)), a = 2.0 * M_PI * u2;
    out[2 * i] = r * std::cos(a);
    out[2 * i + 1] = r * std::sin(a);
  }
}

void oracle_hankel(int n, const double* x, double* out, long long cnt) {
  for (ll i = 0; i < cnt; i++) {
    cd h = hankel1(n, x[i]);
    out[2 * i] = h.real();
    out[2 * i + 1] = h.imag();
  }
}

void oracle_fejer(int n, double* nodes, double* weights) {
  std::vector<double> a, b;
  fejer1(n, a, b);
  std::copy(a.begin(), a.end(), nodes);
  std::copy(b.begin(), b.end(), weights);
}

void oracle_rp_cov(int which, const double* t, long long n, double alpha, double p, double* xi,
                   double* dxi) {
  for (ll i = 0; i < n; i++) {
    if (which == 0) {
      xi[i] = xi_cov_center(t[i], alpha, p);
      dxi[i] = dxi_cov_center(t[i], alpha, p);
    } else if (which == -1) {
      xi[i] = xi_cov_left(t[i], p);
      dxi[i] = dxi_cov_left(t[i], p);
    } else {
      xi[i] = xi_cov_right(t[i], p);
      dxi[i] = dxi_cov_right(t[i], p);
    }
  }
}

// Curve description shared with the product ABI: type 0 circle (cx, cy, R),
// 1 kite (A, B, C, D), 2 star (cx, cy, R, nmodes, a[0..], b[0..]).
static Curve make_curve(int type, const double* prm, int nprm) {
  Curve c;
  c.type = type;
  if (type == 0) {
    c.cx = prm[0];
    c.cy = prm[1];
    c.R = prm[2];
  } else if (type == 1) {
    c.A = prm[0];
    c.B = prm[1];
    c.C = prm[2];
    c.D = prm[3];
  } else {
    c.cx = prm[0];
    c.cy = prm[1];
    c.R = prm[2];
    const int nm = (int)prm[3];
    if (nprm < 4 + 2 * nm) throw std::invalid_argument("star parameters too short");
    c.am.assign(prm + 4, prm + 4 + nm);
    c.bm.assign(prm + 4 + nm, prm + 4 + 2 * nm);
  }
  return c;
}

// Curve::xt / yt / xpt / ypt and Curve::diff at n parameters (Geometry/Curve.hpp:97-146).
int oracle_curve_eval(int type, const double* prm, int nprm, const double* t, long long n, double* x, double* y,
                      double* xp, double* yp) {
  try {
    const Curve c = make_curve(type, prm, nprm);
    for (ll i = 0; i < n; i++) {
      x[i] = c.xt(t[i]);
      y[i] = c.yt(t[i]);
      xp[i] = c.xpt(t[i]);
      yp[i] = c.ypt(t[i]);
    }
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
  return 0;
}
int oracle_curve_diff(int type, const double* prm, int nprm, const double* a, const double* d, long long n,
                      double* dx, double* dy) {
  try {
    const Curve c = make_curve(type, prm, nprm);
    for (ll i = 0; i < n; i++) c.diff(a[i], d[i], dx[i], dy[i]);
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
  return 0;
}

int oracle_patch_lims(int type, const double* prm, int nprm, double wl_per_patch, double k,
                      double* out, int cap) {
  try {
    auto lims = compute_patch_lims(make_curve(type, prm, nprm), wl_per_patch, k);
    if ((int)lims.size() > cap) return -(int)lims.size();
    std::copy(lims.begin(), lims.end(), out);
    return (int)lims.size();
  } catch (std::exception& e) {
    g_err = e.what();
    return 0;
  }
}

struct OracleFM {
  std::unique_ptr<ForwardMap> fm;
};

// lims: npatches+1 parameter limits of one closed curve.
int oracle_fm_create(int type, const double* prm, int nprm, const double* lims, int npatches,
                     int Np, int F, double k, double eta, double delta, double p, int Ps, int Pang,
                     int nlev, void** out) {
  try {
    auto* o = new OracleFM;
    std::vector<Curve> cv{make_curve(type, prm, nprm)};
    std::vector<std::vector<double>> L{std::vector<double>(lims, lims + npatches + 1)};
    o->fm.reset(new ForwardMap(cv, L, Np, F, k, eta, delta, p, Ps, Pang, nlev));
    o->fm->setup();
    *out = o;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

void oracle_fm_destroy(void* h) { delete (OracleFM*)h; }

long long oracle_fm_size(void* h) { return (ll)((OracleFM*)h)->fm->xs.size(); }

void oracle_fm_points(void* h, double* x, double* y, double* nx, double* ny, double* w) {
  auto& f = *((OracleFM*)h)->fm;
  std::copy(f.xs.begin(), f.xs.end(), x);
  std::copy(f.ys.begin(), f.ys.end(), y);
  std::copy(f.nxs.begin(), f.nxs.end(), nx);
  std::copy(f.nys.begin(), f.nys.end(), ny);
  std::vector<cd> one(f.xs.size(), 1.0);
  f.intensities(one);
  for (size_t i = 0; i < one.size(); i++) w[i] = one[i].real();
}

// Near-singular structure: per patch the number of near-singular points, then
// the flat (point index, t) lists.
long long oracle_fm_ns(void* h, int* counts, long long* idx, double* ts) {
  auto& f = *((OracleFM*)h)->fm;
  ll off = 0;
  for (size_t p = 0; p < f.patches.size(); p++) {
    counts[p] = (int)f.patches[p].ns_idx.size();
    for (size_t j = 0; j < f.patches[p].ns_idx.size(); j++) {
      if (idx) idx[off] = f.patches[p].ns_idx[j];
      if (ts) ts[off] = f.patches[p].ns_t[j];
      off++;
    }
  }
  return off;
}

// Precomputation matrix of one patch, Np x (Np + n_ns) complex, column-major.
void oracle_fm_precomp(void* h, int patch, double* out) {
  auto& P = ((OracleFM*)h)->fm->patches[patch];
  std::memcpy(out, P.precomp.data(), P.precomp.size() * sizeof(cd));
}

int oracle_fm_apply(void* h, const double* xin, double* yout, int accelerated) {
  try {
    auto& f = *((OracleFM*)h)->fm;
    const ll N = (ll)f.xs.size();
    std::vector<cd> x(N);
    std::memcpy(x.data(), xin, N * sizeof(cd));
    std::vector<cd> y = accelerated ? f.apply(x) : f.apply_unacc(x);
    std::memcpy(yout, y.data(), N * sizeof(cd));
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

void oracle_fm_stats(void* h, long long* out) {
  auto& f = *((OracleFM*)h)->fm;
  out[0] = f.last.near;
  out[1] = f.last.leaf;
  out[2] = f.last.prop;
  out[3] = f.last.interp;
  out[4] = f.boxes->st_.cones;
}

// Tree tables for cross-checking the product's flattened tables: Morton
// permutation (sorting_), then per level 2..D the relevant-cone list
// (box Morton code, non-relevant cone id) in canonical order.
void oracle_fm_sorting(void* h, long long* sorting) {
  auto& b = *((OracleFM*)h)->fm->boxes;
  std::copy(b.sorting_.begin(), b.sorting_.end(), sorting);
}
long long oracle_fm_level_cones(void* h, int level, long long* box, long long* nonrel) {
  auto& L = *((OracleFM*)h)->fm->boxes->levels_[level];
  if (box) std::copy(L.rel2box.begin(), L.rel2box.end(), box);
  if (nonrel) std::copy(L.rel2nonrel.begin(), L.rel2nonrel.end(), nonrel);
  return (ll)L.rel2box.size();
}

int oracle_fm_propagate(void* h, const double* tx, const double* ty, long long nt,
                        const double* dens, double* out) {
  auto& f = *((OracleFM*)h)->fm;
  const ll N = (ll)f.xs.size();
  std::vector<cd> d(N);
  std::memcpy(d.data(), dens, N * sizeof(cd));
  auto r = f.propagate(std::vector<double>(tx, tx + nt), std::vector<double>(ty, ty + nt), d);
  std::memcpy(out, r.data(), nt * sizeof(cd));
  return 0;
}

int oracle_fm_gmres(void* h, const double* rhs, double* x, int restart, double tol, int maxit,
                    int* iters, double* err, double* hist, int hist_cap) {
  try {
    auto& f = *((OracleFM*)h)->fm;
    const ll N = (ll)f.xs.size();
    std::vector<cd> b(N), xv(N);
    std::memcpy(b.data(), rhs, N * sizeof(cd));
    std::memcpy(xv.data(), x, N * sizeof(cd));
    std::vector<double> hh;
    bool ok = gmres([&](const std::vector<cd>& a, std::vector<cd>& o) { o = f.apply(a); }, b, xv,
                    restart, tol, maxit, *iters, *err, &hh);
    for (int i = 0; i < (int)hh.size() && i < hist_cap; i++) hist[i] = hh[i];
    std::memcpy(x, xv.data(), N * sizeof(cd));
    return ok ? 0 : 2;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// GMRES on a dense column-major complex matrix (pins the solver on small
// systems against a direct solve).
int oracle_gmres_dense(const double* Am, long long n, const double* rhs, double* x, int restart,
                       double tol, int maxit, int* iters, double* err) {
  const cd* A = (const cd*)Am;
  std::vector<cd> b((const cd*)rhs, (const cd*)rhs + n), xv((cd*)x, (cd*)x + n);
  bool ok = gmres(
      [&](const std::vector<cd>& a, std::vector<cd>& o) {
        o.assign(n, 0.0);
        for (ll j = 0; j < n; j++)
          for (ll i = 0; i < n; i++) o[i] += A[j * n + i] * a[j];
      },
      b, xv, restart, tol, maxit, *iters, *err, nullptr);
  std::memcpy(x, xv.data(), n * sizeof(cd));
  return ok ? 0 : 2;
}

// Standalone BoxTree::Solve on arbitrary points (IFGF_Source/main.cpp:178-239 shape).
int oracle_boxtree_solve(const double* x, const double* y, const double* nx, const double* ny,
                         long long N, int nlev, double k, int F, double eta, int Ps, int Pang,
                         const double* dens, double* out, long long* stats) {
  try {
    BoxTree b(std::vector<double>(x, x + N), std::vector<double>(y, y + N),
              std::vector<double>(nx, nx + N), std::vector<double>(ny, ny + N), nlev, k, F, eta,
              Ps, Pang);
    std::vector<cd> d((const cd*)dens, (const cd*)dens + N);
    b.Solve(true, d, nullptr);
    std::memcpy(out, d.data(), N * sizeof(cd));
    if (stats) {
      stats[0] = b.st_.near;
      stats[1] = b.st_.leaf;
      stats[2] = b.st_.prop;
      stats[3] = b.st_.interp;
      stats[4] = b.st_.cones;
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// BoxTree::Solve with a separate target set (no reference counterpart; the
// IFGF-accelerated form of propagate_unacc, ForwardMap.tpp:702-727, taking the
// intensities): out[t] = sum_s G(target t, source s) dens[s], near field
// included, zero for coincident points.
int oracle_field_solve(const double* x, const double* y, const double* nx, const double* ny, long long N,
                       const double* tx, const double* ty, long long M, int nlev, double k, int F, double eta,
                       int Ps, int Pang, const double* dens, double* out, long long* stats) {
  try {
    const std::vector<double> vtx(tx, tx + M), vty(ty, ty + M);
    BoxTree b(std::vector<double>(x, x + N), std::vector<double>(y, y + N),
              std::vector<double>(nx, nx + N), std::vector<double>(ny, ny + N), nlev, k, F, eta,
              Ps, Pang, &vtx, &vty);
    std::vector<cd> d((const cd*)dens, (const cd*)dens + N);
    b.Solve(true, d, nullptr);
    std::memcpy(out, d.data(), M * sizeof(cd));
    if (stats) {
      stats[0] = b.st_.near;
      stats[1] = b.st_.leaf;
      stats[2] = b.st_.prop;
      stats[3] = b.st_.interp;
      stats[4] = b.st_.cones;
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Direct O(N^2) sum skipping only i == j (IFGF_Source/main.cpp:104-148), on a
// subset of targets.
void oracle_direct(const double* x, const double* y, const double* nx, const double* ny,
                   long long N, double k, int F, double eta, const double* dens,
                   const long long* targets, long long nt, double* out) {
  const cd* d = (const cd*)dens;
  for (ll t = 0; t < nt; t++) {
    const ll i = targets[t];
    cd acc = 0.0;
    for (ll s = 0; s < N; s++) {
      if (s == i) continue;
      acc += kernel(F, x[i], y[i], x[s], y[s], nx[s], ny[s], k, eta) * d[s];
    }
    out[2 * t] = acc.real();
    out[2 * t + 1] = acc.imag();
  }
}

}  // extern "C"
