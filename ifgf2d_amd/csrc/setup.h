// Host-side setup of the IFGF forward map: geometry, near-singular detection,
// the rectangular-polar (RP) singular/near-singular correction, and the
// flattened Morton tree / cone tables that the gfx950 kernels read.
//
// Every discrete decision (Morton box of a point, relevant boxes and cones,
// the cone of every (target, cousin) pair, the child cone of every parent
// interpolation node) is made here, once per geometry, with the reference's
// formulas; the kernels only do arithmetic on the resulting index tables.
#pragma once

#include <string>
#include <vector>

#include "rp.h"
#include "zmath.h"

namespace ifgf {

bool make_curve(int type, const double* prm, int nprm, Curve& c, std::string& err);

// Nystrom discretisation in patch order (ForwardMap.tpp:37-222).
struct Geometry {
  int Np = 0, npatches = 0;
  long long N = 0;
  std::vector<double> t1, t2;                   // per patch
  std::vector<double> x, y, nx, ny, w, t;       // per point; w = intensity factor
  std::vector<double> fejer_nodes, fejer_weights;
  // per patch: near-singular point indices and their closest parameter
  std::vector<std::vector<long long>> ns_idx;
  std::vector<std::vector<double>> ns_t;
};
// Near-singular points: the points of the two adjacent patches (e = 0: the
// previous, 1: the next, cyclically -- the neighbour estimate of
// ForwardMap.tpp:164-218 for one closed curve) closer than delta to the patch,
// with their closest parameter (near_singular_param). ns_search (optional)
// fills g.ns_idx / g.ns_t in that order instead of the host loop (the device
// search).
IFGF_HD inline int near_singular_candidate_patch(int p, int e, int npatches) {
  return e == 0 ? (p - 1 + npatches) % npatches : (p + 1) % npatches;
}
using NsSearch = bool (*)(const Curve& c, double delta, Geometry& g, std::string& err);
bool build_geometry(const Curve& c, const double* lims, int npatches, int Np, double delta,
                    Geometry& g, std::string& err, NsSearch ns_search = nullptr);

// Folded correction W_p = M_p^T T_Np in target-major CSR: for point i (patch
// order), entries e in [ptr[i], ptr[i+1]) add sum_k w[e*Np+k] * rho[patch[e]*Np+k].
struct Correction {
  std::vector<long long> ptr;
  std::vector<int> patch;
  std::vector<zd> w;
};
bool build_correction(const Geometry& g, const Curve& c, int F, double k, double eta, double p,
                      Correction& out, std::string& err);
// The RP columns of every patch in the order the correction uses them: the
// patch's own Np nodes (tsing = ttgt = the node's parameter), then its
// near-singular points (tsing = closest parameter, ttgt = the point's).
struct RpColumn {
  int patch;
  double tsing, ttgt;
};
void rp_columns(const Geometry& g, std::vector<RpColumn>& cols, std::vector<long long>& patch_col_off);
// Chebyshev-coefficient map folded into the columns: Tfold[n * Np + k] =
// (alpha_n / Np) T_n(x_k) (Cheb1D::interp_1d<Np>, Chebyshev1d.hpp:134-163).
std::vector<double> rp_fold_table(const Geometry& g);
// Target-major CSR of the folded columns W (patch_col_off[p] * Np + col * Np + k).
void assemble_correction(const Geometry& g, const std::vector<zd>& W, const std::vector<long long>& patch_col_off,
                         Correction& out);
// 40-node Fejer rule of the RP integrals (num_ns, ForwardMap.hpp:70).
void rp_rule(std::vector<double>& fn, std::vector<double>& fw);

struct LevelInfo {
  int level = 0;
  double min_x = 0, min_y = 0, bs = 0, h = 0;
  int nce = 0, nrad = 0, radial_off = 0;
  int nbox = 0, box_off = 0;
  int ncone = 0, cone_off = 0;
  int nslot = 0;          // non-relevant cone slots per box: (2 nce + 1) * nrad
  long long map_off = 0;  // dense (box, nonrel) -> local relevant cone
  long long proto_off = -1;  // protobox of this (child) level, indexed by parent nonrel
};

struct Tree {
  int nlev = 0, D = 0, Ps = 0, Pang = 0;
  long long N = 0;
  double k = 0;
  std::vector<int> perm, inv;                  // sorted -> original, original -> sorted
  std::vector<double> sx, sy, snx, sny;        // Morton-sorted points
  std::vector<LevelInfo> lv;                   // indexed by level, valid for 2..D
  std::vector<double> radial;                  // all levels' radial interval ends
  std::vector<long long> box_morton;           // global box id (levels 2..D)
  std::vector<double> box_cx, box_cy;
  std::vector<int> box_child;                  // 4 per global box: local child index or -1
  std::vector<int> leaf_begin;                 // leaf boxes: nleaf + 1 point offsets
  std::vector<int> leaf_of_point;              // sorted point -> leaf box (local)
  std::vector<int> nb_ptr, nb_list;            // leaf neighbours (local leaf ids)
  std::vector<int> cone_box, cone_posr, cone_post;   // global cone id -> local box, posr, postheta
  std::vector<unsigned char> cone_level;
  std::vector<int> dense_map;                  // see LevelInfo::map_off
  std::vector<int> proto;                      // see LevelInfo::proto_off, [q][4][P]
  std::vector<long long> it_ptr;               // interpolation entries per sorted target
  std::vector<int> it_cone;                    // global cone id per entry
  std::vector<int> point_box;                  // sorted point x level -> local box (N x (D+1))
  long long n_leaf = 0, n_prop = 0, n_interp = 0, n_cones = 0;
  // Targets. The forward map's targets are its sorted sources (sep = false,
  // M = N, the arrays below empty). A field tree (sep = true) holds M
  // separate targets sorted by leaf Morton code (stable), the leaf group of
  // every sorted target (targets sharing a leaf box) and, per group, the
  // source leaf boxes of its 3x3 neighbourhood in ascending Morton order.
  // it_ptr / it_cone then run over the sorted targets.
  bool sep = false;
  long long M = 0;
  std::vector<int> tperm, tinv;                // sorted target -> original, original -> sorted
  std::vector<double> tx, ty;                  // sorted targets
  std::vector<int> tleaf, tnb_ptr, tnb_list;
  long long n_near_t = 0;                      // (target, source) pairs of the field's near field
  // device cone builder: interpolation entries and protobox nodes within a
  // relative 1e-9 of a cone boundary, decided on the host
  long long near_boundary_entries = 0, near_boundary_nodes = 0;
};
// Inputs of the relevant-cone step (interpolation entries, protoboxes and the
// numbering of the relevant cones) when build_tree hands it to a builder of
// its own: the sorted targets, their group at every level, each group's
// cousin boxes and the Chebyshev nodes.
struct ConeInput {
  const double* tx = nullptr;  // sorted targets
  const double* ty = nullptr;
  long long nt = 0;
  const int* pbox = nullptr;      // nt x (D + 1): target group per level
  std::vector<long long> grp_off;  // first global group id per level (D + 2 entries)
  std::vector<int> cz_ptr, cz_list;  // cousin boxes (local ids, ascending) per global group
  std::vector<double> xs, xa;        // Chebyshev nodes in s and theta
};
// A builder fills t.dense_map, t.proto, t.cone_box / posr / post / level,
// lv[l].cone_off / ncone, t.n_cones, t.n_prop and t.it_cone exactly as the
// host loop does (lv[l].map_off and proto_off are set before the call).
using ConeBuilder = bool (*)(const ConeInput& in, Tree& t, std::string& err);
// The builder on the GPU (cones.hip).
bool relevant_cones_device(const ConeInput& in, Tree& t, std::string& err);

// Inputs of the grouping step (each target's box at every level, the target
// groups, their cousin boxes, the leaf neighbourhoods and the interpolation
// entry counts) when build_tree hands it to a builder of its own.
struct GroupJob {
  const double* tx = nullptr;  // sorted targets
  const double* ty = nullptr;
  long long nt = 0;
  bool sep = false;             // separate targets (a field tree)
  std::vector<long long> nbps;  // boxes per side, per level
};
// A builder fills pbox (nt x (D + 1), -1 below level 2), the cousin CSR of
// cin (grp_off, cz_ptr, cz_list), t.it_ptr and the leaf neighbourhoods
// (t.nb_ptr / nb_list of the source leaf boxes for the forward map;
// t.tnb_ptr / tnb_list and t.tleaf for a field) exactly as the host loops do.
using GroupBuilder = bool (*)(const GroupJob& job, Tree& t, std::vector<int>& pbox, ConeInput& cin,
                              std::string& err);
// The builder on the GPU (cones.hip).
bool tree_groups_device(const GroupJob& job, Tree& t, std::vector<int>& pbox, ConeInput& cin, std::string& err);

// perm_in (optional): the Morton permutation computed elsewhere (the device
// sort); it is checked to be a permutation that orders the leaf codes with
// ties in input order, i.e. exactly the host's stable sort.
// tx, ty, M (optional): a separate target set (a field tree). The bounding
// box then covers sources and targets; relevant cones are those the targets'
// interpolations and the parents need, exactly as for the forward map with
// the targets in place of the sources.
// cones (optional): builds the relevant cones in place of the host loop.
// groups (optional, with cones): builds the target groups, cousin and
// neighbour lists and entry counts in place of the host loops.
bool build_tree(const double* x, const double* y, const double* nx, const double* ny, long long N,
                int nlev, double k, int Ps, int Pang, Tree& t, std::string& err,
                const std::vector<int>* perm_in = nullptr, const double* tx = nullptr,
                const double* ty = nullptr, long long M = 0, ConeBuilder cones = nullptr,
                GroupBuilder groups = nullptr);

// Near-field exclusion lists (init_sort_sing_point, ForwardMap.tpp:567-607) in
// sorted order: own patch + patches the point is near-singular to, padded
// with -1, kMaxExcl per point; patch id per sorted point. Also checks that all
// excluded pairs are leaf-level neighbours (BoxTree.h:1079-1161).
constexpr int kMaxExcl = 4;
bool build_exclusions(const Geometry& g, const Tree& t, std::vector<int>& patch_of_sorted,
                      std::vector<int>& excl, long long& n_near, std::string& err);

}  // namespace ifgf
