/*
 * ifgf2d.h -- C ABI of the MI355X IFGF forward map (libifgf2d.so).
 *
 * The drop-in seam of manuelarturosantana/IFGF2D is the operator the GMRES
 * driver applies once per iteration:
 *   LinearOperator(rows, cols, mapping)          ForwardMap/LinearOperator.hpp:27-60
 *   mapping = [&](x, out){ compute_Ax_acc(x, out, boxes); }   ForwardMap.tpp:673-680
 *   out = mat * v                                 ForwardMap/opGMRES.hpp:102,152,250
 * Every entry point below replaces one reference interface (cited per
 * function). Plain pointers and sizes only; complex vectors are interleaved
 * (re, im) doubles, the std::complex<double> / Eigen::VectorXcd layout, in the
 * reference's patch order (patch-major, Np nodes per patch). No exceptions
 * cross this boundary: every function returns a status code (IFGF_OK = 0) and
 * ifgf_last_error() / ifgf_plan_error() return the message.
 */
#ifndef IFGF2D_H
#define IFGF2D_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define IFGF_OK 0
#define IFGF_ERR_ARG 1          /* invalid argument */
#define IFGF_ERR_GEOMETRY 2     /* degenerate geometry / tree invariant violated */
#define IFGF_ERR_LEVELS 3       /* nlevels < 3 (BoxTree.h:1024-1025) */
#define IFGF_ERR_HIP 4          /* HIP runtime error */
#define IFGF_ERR_NO_DEVICE 5    /* plan was built host-only, or no GPU */
#define IFGF_ERR_NOT_CONVERGED 6 /* GMRES reached maxit (Eigen::NoConvergence) */

/* formulations (GreenFunctions.hpp:15-20) */
#define IFGF_SINGLE_LAYER 0
#define IFGF_DOUBLE_LAYER 1
#define IFGF_COMBINED 2

/* precision modes */
#define IFGF_FP64 0   /* everything in FP64 */
#define IFGF_MIXED 1  /* FP32 leaf operator, cone coefficients, propagation and interpolation;
                         FP64 Hankel, centre factors (evaluated in FP64, applied rounded to
                         FP32), near field, correction and final accumulation (8-column
                         blocks: the near field on FP32 matrix cores, its sums widened to
                         FP64). The plan also keeps the FP64 operator (IFGF_APPLY_FP64).
                         Needs stored operators. */

/* curve types (Geometry/Curve.hpp:106-146; star has no reference class) */
#define IFGF_CURVE_CIRCLE 0 /* params: cx, cy, R */
#define IFGF_CURVE_KITE 1   /* params: A, B, C, D: x = A cos t + B cos 2t + C, y = D sin t */
#define IFGF_CURVE_STAR 2   /* params: cx, cy, R, nmodes, a[nmodes], b[nmodes] */

/* plan flags */
#define IFGF_PLAN_HOST_ONLY 1  /* build every host table, allocate nothing on the GPU */
#define IFGF_PLAN_ON_THE_FLY 2 /* re-evaluate the near-field and leaf kernels in every
                                  application instead of storing them (default: the
                                  geometry-only near-field values and the folded leaf
                                  operator are evaluated once at plan creation and
                                  streamed from HBM by every application) */
#define IFGF_PLAN_AUTOTUNE 4   /* time the candidate lane counts of the stored-path
                                  kernels on this plan and keep the fastest */
#define IFGF_PLAN_DEVICE_SORT 8 /* Morton box assignment and the stable sort of the points
                                   on the GPU (ifgf_morton_sort) instead of the host; the
                                   permutation is checked to equal the host's stable sort */
#define IFGF_PLAN_OVERLAP 16   /* run the stored near field and the identity / correction half of
                                  the final sum on a second stream beside the level pushes
                                  (IFGF_PLAN_AUTOTUNE also tries it) */
#define IFGF_PLAN_DEVICE_PRECOMP 32 /* evaluate the rectangular-polar singular / near-singular
                                       correction (ForwardMap.tpp:287-423) on the GPU, one thread
                                       per column, instead of on the host (Np <= 16) */
#define IFGF_PLAN_DEVICE_GEOMETRY 64 /* the near-singular search of the discretisation
                                        (ForwardMap.tpp:224-285) on the GPU, one thread per
                                        candidate point, instead of on the host */
#define IFGF_PLAN_DEVICE_TREE 128 /* the tree's target groups, cousin and neighbour lists
                                     (Level.h:145-161, 261-289, 378-404) and relevant cone segments
                                     (interpolation entries, protoboxes, numbering;
                                     BoxTree.h:299-583) on the GPU instead of the host; same tables
                                     as the host's loops */

#define IFGF_PLAN_NO_MFMA 256 /* blocks of 8 columns (nrhs_max >= 8) use the VALU batched near-field and
                                 leaf kernels instead of the matrix-core (MFMA) ones */

/* apply flags */
#define IFGF_PTR_DEVICE 1     /* x / y are device pointers; the call is asynchronous on stream */
#define IFGF_APPLY_FP64 2     /* on an IFGF_MIXED plan: apply the all-FP64 operator instead of the
                                 mixed-precision one, in the same column blocks (no effect on an
                                 IFGF_FP64 plan) */

typedef struct ifgf_plan ifgf_plan;
typedef struct ifgf_group ifgf_group;

/* Discretised closed curve: equal-weight Fejer first-rule nodes on each patch
 * [patch_lims[p], patch_lims[p+1]] (ForwardMap ctor, ForwardMap.tpp:15-222). */
typedef struct {
  int curve_type;
  const double* curve_params;
  int ncurve_params;
  const double* patch_lims; /* npatches + 1 increasing parameters spanning [0, 2 pi] */
  int npatches;
  int Np;        /* nodes per patch (template Np of ForwardMap) */
  double delta;  /* near-singular distance (ForwardMap::delta_) */
  double rp_p;   /* rectangular-polar change-of-variables strength (ForwardMap::p_) */
} ifgf_surface;

typedef struct {
  int formulation;
  int precision;
  double k;       /* wavenumber */
  double eta;     /* combined-field coupling */
  int Ps, Pang;   /* Chebyshev nodes per cone in s and theta (BoxTree<Ps,Pang>) */
  int nlevels;    /* tree levels, >= 3 */
  int nrhs_max;   /* largest nrhs passed to ifgf_apply (>= 16: a second lane of 8-column buffers;
                     >= 24: a third, which IFGF_PLAN_AUTOTUNE drops where two time faster) */
  int flags;      /* IFGF_PLAN_HOST_ONLY | IFGF_PLAN_ON_THE_FLY | IFGF_PLAN_AUTOTUNE */
} ifgf_params;

/* Replaces ForwardMap::ForwardMap (ForwardMap.tpp:15-35) followed by
 * ForwardMap::precomps_and_setup (ForwardMap.tpp:542-565): geometry, RP
 * precomputations, BoxTree::Reinitialize, the neighbourhood check and
 * init_sort_sing_point. Errors: IFGF_ERR_LEVELS, IFGF_ERR_GEOMETRY (the
 * std::logic_error cases of Level.h:153-157 and BoxTree.h:321,571,1117,1152),
 * IFGF_ERR_ARG, IFGF_ERR_HIP. */
int ifgf_plan_create(const ifgf_surface* surface, const ifgf_params* params, ifgf_plan** plan);

/* Releases the plan (BoxTree / ForwardMap destructors). */
void ifgf_plan_destroy(ifgf_plan* plan);

/* Message of the last failing call on this thread / on this plan. */
const char* ifgf_last_error(void);
const char* ifgf_plan_error(const ifgf_plan* plan);

/* LinearOperator::rows() / cols() (LinearOperator.hpp:48-49). */
long long ifgf_plan_size(const ifgf_plan* plan);

/* ForwardMap::xs_, ys_, nxs_, nys_ (ForwardMap.hpp:72-75) and the intensity
 * factor of compute_intensities (ForwardMap.tpp:425-435); any may be NULL. */
int ifgf_plan_points(const ifgf_plan* plan, double* x, double* y, double* nx, double* ny, double* w);

/* Work counts of one application: [0] near-field pairs, [1] leaf kernel
 * evaluations, [2] propagation interpolations, [3] target interpolations,
 * [4] relevant cones (all levels), [5] correction entries, [6] nlevels,
 * [7] N, [8] leaf-level cones, [9] 1 if the near-field values and leaf
 * operator are stored, [10] stored near-field entries (excluded pairs
 * included as zeros), [11] bytes of the stored operators, [12] and [13] lanes
 * per target / leaf-cone node of the near-field and leaf kernels, [14]
 * evaluations per block of the per-level propagation/interpolation kernel,
 * [15] lanes per target of the correction kernel, [16] the near field on a second
 * stream beside the level pushes: 0 no, 1 forked after the leaf stage, 2 after the
 * first level push (the first half of the correction kernel runs there too),
 * [17] 1 on a mixed-precision plan,
 * [18] stored propagation-geometry records (one per level, parent cone
 * direction, child position and node), [19] stored inverse node factors. */
int ifgf_plan_stats(const ifgf_plan* plan, long long out[20]);

/* Replaces BoxTree::SortBox (BoxTree.h:73-106) with ComputeBB's bounding box
 * (BoxTree.h:179-202) and Level::Point2Morton (Level.h:145-161), on the GPU:
 * the leaf boxes of an nlevels tree (leaf grid 2^(nlevels-1) per side,
 * box index floor((x - min) / box size), x in the low bit of the interleaved
 * code) and the stable sort of the points by code. sorting[i] = input index
 * of the i-th point in Morton order (BoxTree::sorting_); codes (may be NULL)
 * = leaf code of every input point. Host pointers. nlevels in [3, 16].
 * Errors: IFGF_ERR_LEVELS, IFGF_ERR_GEOMETRY (a point outside the leaf grid). */
int ifgf_morton_sort(const double* x, const double* y, long long n, int nlevels, long long* sorting,
                     long long* codes);

/* BoxTree::getSorting() (BoxTree.h:1163-1165): Morton position -> patch-order index. */
int ifgf_plan_sorting(const ifgf_plan* plan, long long* sorting);

/* Level::relconesegment2cocenteredmortonboxid_ / relconesegment2nonrelconesegment_
 * (Level.h:41-43) of one level (2..nlevels-1), in the plan's canonical order
 * (box Morton ascending, non-relevant cone ascending). Returns the count, or
 * -1 on error; arrays may be NULL to query the count. */
long long ifgf_plan_level_cones(const ifgf_plan* plan, int level, long long* box_morton,
                                long long* nonrel);

/* Folded correction rows (compute_sing_near_sing_interactions,
 * ForwardMap.tpp:481-515, with Cheb1D::interp_1d folded in): ptr has N + 1
 * entries; entry e adds sum_k w[e][k] x[patch[e]*Np + k] to target i. Any
 * array may be NULL; returns the entry count, or -1 on error. */
long long ifgf_plan_correction(const ifgf_plan* plan, long long* ptr, int* patch, double* w);

/* The forward map: replaces ForwardMap::compute_Ax_acc(const VectorXcd&, VectorXcd&,
 * BoxTree&) (ForwardMap.tpp:648-655), i.e. the mapping LinearOperator::operator*
 * calls (LinearOperator.cpp:14-23). y = (1/2) x [double layer / combined only]
 * + correction(x) + IFGF(intensities(x)). x, y: nrhs columns of n complex
 * values, column stride ld (>= n). Host pointers: synchronous. With
 * IFGF_PTR_DEVICE: device pointers, asynchronous on stream (hipStream_t, NULL
 * = the plan's stream). Applies on one plan are serialised; the plan is not
 * reentrant (like BoxTree's mutable buffers, BoxTree.h:48-51). */
int ifgf_apply(ifgf_plan* plan, const double* x, double* y, int nrhs, long long ld, int flags,
               void* stream);

/* ifgf_apply with IFGF_PTR_DEVICE as one HIP graph (no reference counterpart;
 * for callers that replay the same application many times, e.g. a solver
 * loop over fixed Krylov buffers). The first call with a given (x, y, nrhs,
 * ld, IFGF_APPLY_FP64) captures the application's launches on the plan's
 * stream into a graph (one single-column application per column: the batched
 * column blocks and their lanes are not captured); every call launches that
 * graph on stream (NULL = the
 * plan's stream), asynchronously. x / y are device pointers whose contents
 * may change between calls; a call with other pointers recaptures. Fails
 * while stage profiling is on. */
int ifgf_apply_graph(ifgf_plan* plan, const double* x, double* y, int nrhs, long long ld, int flags,
                     void* stream);

/* Several obstacles' forward maps as one launch (no reference counterpart;
 * the reference handles one ForwardMap at a time): y[i] = A_i x[i] for nplans
 * distinct plans with stored operators (device pointers, n_i complex values
 * each, one column). Every stage of the application runs as ONE kernel launch
 * over all plans of a class (plans of equal (Ps, Pang) and precision; the
 * formulation may differ): gather, near field, leaf operator, one push
 * launch per tree depth, finalize; each workgroup finds its plan from a table
 * of per-plan block offsets, so every plan computes exactly what its own
 * application computes (bitwise). The launches are captured into one HIP
 * graph; every ifgf_group_launch runs it on stream (NULL = the group's own
 * stream), asynchronously. x[i] / y[i] contents may change between launches,
 * the pointers may not. Plans must outlive the group and are not applied
 * elsewhere while a launch runs. */
int ifgf_group_create(ifgf_plan* const* plans, int nplans, const double* const* x, double* const* y,
                      ifgf_group** group);
int ifgf_group_launch(ifgf_group* group, void* stream);
int ifgf_group_synchronize(ifgf_group* group);  /* waits for the group's own stream */
/* Cone-order autotuner (no reference counterpart: the reference fixes Ps
 * and Pang at compile time, ForwardMap.hpp:49, BoxTree.h:24): builds the
 * surface's plan for (Ps, Pang) = (3, 5), (5, 5), (5, 7), (7, 7), (9, 9),
 * (11, 11) in that order of cost and returns the first whose application of
 * a seeded random density is within tol (relative 2-norm) of the
 * unaccelerated operator (ifgf_apply_direct) at up to 256 targets spread
 * over the points; params gives everything else. err: that order's error.
 * IFGF_ERR_NOT_CONVERGED (with the most accurate order) when none reaches tol. */
int ifgf_autotune_cone_order(const ifgf_surface* surface, const ifgf_params* params, double tol, int* Ps, int* Pang,
                             double* err);

/* out[4]: plans, classes (plans of equal (Ps, Pang) and precision share every
 * launch), kernel launches per group application, workgroups per application. */
int ifgf_group_stats(const ifgf_group* group, long long out[4]);
void ifgf_group_destroy(ifgf_group* group);

/* Unaccelerated operator on a target subset: replaces compute_Ax_unacc
 * (ForwardMap.tpp:438-540) evaluated at targets[0..nt). Host pointers. */
int ifgf_apply_direct(ifgf_plan* plan, const double* x, const long long* targets, long long nt,
                      double* out);

/* Field at off-surface points: replaces ForwardMap::propagate_unacc
 * (ForwardMap.tpp:702-727) (direct sum on the GPU). Host pointers. */
int ifgf_propagate(ifgf_plan* plan, const double* tx, const double* ty, long long nt,
                   const double* density, double* out);

/* ifgf_propagate on device pointers, allocating nothing: no reference
 * counterpart; for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation). tx, ty: nt
 * target coordinates; density: the plan's N complex values in patch order;
 * out: nt complex values; all device pointers. Asynchronous on stream (NULL =
 * the plan's stream), with the plan's device current. The intensity-weighted
 * density goes through a fixed device scratch array of the library (16 MiB
 * per device, held until the process ends), so the plan may have at most
 * 1,048,576 points (IFGF_ERR_ARG above that). Every call on one device shares
 * that scratch: calls that overlap on one device (other streams, other
 * threads) race and must be serialised by the caller. Same sum as
 * ifgf_propagate, taken over the points in the plan's Morton order. */
int ifgf_propagate_noalloc(ifgf_plan* plan, const double* tx, const double* ty, long long nt, const double* density,
                           double* out, void* stream);

/* IFGF-accelerated field at separate target points (the acceleration has no
 * reference counterpart: the reference sums the same field directly in
 * ForwardMap::propagate_unacc, ForwardMap.tpp:702-727, which ifgf_propagate
 * restates). out[t] = sum_s G(target t, x_s) rho_s over the plan's N points,
 * rho = density x quadrature weight (compute_intensities,
 * ForwardMap.tpp:425-435), G the plan's formulation kernel (k, eta, source
 * normals); coincident points contribute zero. The targets get an IFGF tree
 * of their own (bounding box of the points and the targets, nlevels levels,
 * Ps x Pang cone interpolation, BoxTree.h:179-583 with the targets in place of
 * the points): each target's 3x3 leaf neighbourhood is summed exactly, the
 * rest through the cones, so the relative error is the IFGF interpolation
 * error (about 2e-3 at 5 x 5, 1e-7 at 11 x 11 on the unit circle at k = 10).
 * All tables are built once here; each ifgf_field_apply streams them.
 * precision: IFGF_FP64 or IFGF_MIXED. Errors: IFGF_ERR_ARG (nt < 1, unknown
 * (Ps, Pang) or precision), IFGF_ERR_LEVELS, IFGF_ERR_GEOMETRY, IFGF_ERR_HIP.
 * The field of a host-only plan holds the host tables only (stats work,
 * ifgf_field_apply returns IFGF_ERR_NO_DEVICE); the field of a plan created
 * with IFGF_PLAN_AUTOTUNE is tuned the same way. The field keeps its own copy
 * of the points; the plan may be destroyed afterwards. */
typedef struct ifgf_field ifgf_field;
int ifgf_field_create(const ifgf_plan* plan, const double* tx, const double* ty, long long nt, int nlevels,
                      int Ps, int Pang, int precision, ifgf_field** out);
/* density: N complex values in patch order; out: nt complex values in target
 * order. flags: IFGF_PTR_DEVICE (device pointers, stream-ordered on stream or
 * on the field's own stream when stream is NULL; host pointers otherwise,
 * synchronous), IFGF_APPLY_FP64 (all-FP64 operator on an IFGF_MIXED field). */
int ifgf_field_apply(ifgf_field* field, const double* density, double* out, int flags, void* stream);
/* out[8]: targets, near-field pairs, leaf evaluations, propagation entries,
 * interpolation entries, relevant cones, levels, stored-operator bytes. */
int ifgf_field_stats(const ifgf_field* field, long long out[8]);
void ifgf_field_destroy(ifgf_field* field);

/* Restarted GMRES on the GPU with the plan's forward map: replaces
 * ForwardMap::solve's Eigen::opGMRES<LinearOperator> (ForwardMap.tpp:658-697,
 * opGMRES.hpp:68-269). Arnoldi by classical Gram-Schmidt with a second
 * pass when the first leaves less than a tenth of the new vector's norm
 * (FP64). Stops when the recursive residual estimate
 * |w(k)| / ||r0|| < tol (opGMRES.hpp:207) or after maxit iterations.
 * rhs / x are host pointers of n complex values; x holds the initial guess.
 * relres returns the true residual ||b - A x|| / ||b|| at exit, always with
 * the all-FP64 operator (on an IFGF_MIXED plan the iterations use the
 * mixed-precision operator, whose true residual levels off near its FP32
 * rounding; use ifgf_fgmres there). hist (may be NULL) receives up to
 * hist_cap residual estimates. Returns IFGF_OK, IFGF_ERR_NOT_CONVERGED or an
 * error. */
int ifgf_gmres(ifgf_plan* plan, const double* rhs, double* x, int restart, double tol, int maxit,
               int* iters, double* relres, double* hist, int hist_cap);

/* GMRES for nrhs right-hand sides of the plan's operator at once (no
 * reference counterpart; the reference solves one right-hand side per
 * opGMRES call): every system runs ifgf_gmres's algorithm (restart, the
 * estimate |w(k)| / ||r0|| < tol per system), and the systems advance in
 * lockstep so that each iteration applies the operator to all active systems'
 * new basis vectors in one batched call (blocks of 8 columns when the plan's
 * nrhs_max allows) and orthogonalises all of them with one set of kernels.
 * rhs / x: n x nrhs, column-major, host pointers (x holds the initial
 * guesses; an all-zero x is not copied or applied: the first residuals are
 * rhs). iters[s] / relres[s]: iterations and true FP64 relative residual
 * of system s. Returns IFGF_OK when every system converged. restart <= 256. */
int ifgf_gmres_multi(ifgf_plan* plan, const double* rhs, double* x, int nrhs, int restart, double tol,
                     int maxit, int* iters, double* relres);

/* Mixed-precision refinement for many right-hand sides (no reference
 * counterpart): x <- x + d, d the lockstep GMRES(restart) solution of
 * A d = b - A x (ifgf_gmres_multi, x0 = 0, the plan's default operator: mixed
 * precision, with an FP32 Krylov basis, on an IFGF_MIXED plan) to relative tolerance
 * min(0.5, max(inner_floor, tol / e)), e the largest true relative residual of
 * the systems not yet converged; the residuals use the all-FP64 operator, in
 * the same column blocks. A system stops once its true FP64 relative residual
 * is below tol; at most max_outer refinement steps. rhs / x: n x nrhs,
 * column-major, host pointers (x holds the initial guesses; an all-zero x is
 * not copied or applied: the first residuals are rhs). outer: the
 * refinement steps taken; iters[s]: system s's inner Arnoldi steps; relres[s]:
 * its true FP64 relative residual. Returns IFGF_OK when every system
 * converged, else IFGF_ERR_NOT_CONVERGED. */
int ifgf_refine_multi(ifgf_plan* plan, const double* rhs, double* x, int nrhs, int restart, double tol,
                      int maxit, int max_outer, double inner_floor, int* outer, int* iters, double* relres);

/* Flexible GMRES with FP64 refinement (no reference counterpart; the
 * reference solves with opGMRES only, opGMRES.hpp:68-269): outer FGMRES(restart)
 * in FP64 with the all-FP64 operator, right-preconditioned by an inner GMRES
 * cycle (inner_restart iterations at most, x0 = 0) with the plan's default
 * operator (mixed precision on an IFGF_MIXED plan). The inner relative
 * tolerance for outer step j is min(0.5, max(inner_floor, tol / e_j)), e_j the
 * outer relative residual estimate so far. Stops when the outer estimate
 * |w(k)| / ||r0|| < tol or after maxit outer iterations. outer_iters /
 * inner_iters return the FP64 Arnoldi steps and the inner operator
 * applications' Arnoldi steps; relres the true FP64 relative residual. */
int ifgf_fgmres(ifgf_plan* plan, const double* rhs, double* x, int restart, double tol, int maxit,
                int inner_restart, double inner_floor, int* outer_iters, int* inner_iters,
                double* relres, double* hist, int hist_cap);

/* Wall-clock seconds of the plan's last ifgf_gmres / ifgf_fgmres call: [0]
 * workspace set-up (the Krylov spaces, right-hand side, iterate and FGMRES's
 * preconditioned directions are allocated on the first call and kept by the
 * plan, so a warm call spends ~0 here), [1] host time blocked on the device
 * (the Hessenberg columns and norms the host's Givens steps wait for), [2] the
 * inner GMRES cycles (FGMRES; 0 for GMRES), [3] the whole call. No reference
 * counterpart (opGMRES.hpp:68-269 times nothing). */
int ifgf_plan_solver_times(const ifgf_plan* plan, double seconds[4]);

/* ---- learned preconditioner: a 1D Fourier neural operator ------------------
 * No reference counterpart. The reference's GMRES applies a preconditioner
 * object at opGMRES.hpp:103,153,251 (precond.solve), IdentityPreconditioner
 * by default (:274); these entry points put a learned one there. The network
 * (4 Fourier layers, width 64, 32 modes on the boundary parameter t, BF16
 * weights and activations into FP32-accumulating MFMA, FP32 master weights)
 * maps per-node features of a vector v on the curve (geometry, |v|, its local
 * tangential wavenumber) to a complex multiplier field m and returns
 * z = M(v) = m v: the initial density for v = b, a flexible right
 * preconditioner for Krylov vectors. It is trained on the device (Adam) on
 * (right-hand side, density) pairs that the solver produced. One network
 * object is not reentrant (its workspaces are shared by its calls). */
typedef struct ifgf_fno ifgf_fno;
#define IFGF_FNO_GUESS 1        /* ifgf_fno_gmres: initial density M(b) (when x is zero) */
#define IFGF_FNO_PRECONDITION 2 /* ifgf_fno_gmres: flexible right preconditioner z_j = M(v_j) */
#define IFGF_FNO_FIXED_MULTIPLIER 4 /* with IFGF_FNO_PRECONDITION: z_j = m v_j with the multiplier
                                        field m = M(b) / b the network gives the right-hand side */

long long ifgf_fno_num_params(void);
/* seeded uniform initialisation (splitmix64) */
int ifgf_fno_create(unsigned long long seed, ifgf_fno** out);
void ifgf_fno_destroy(ifgf_fno* fno);
const char* ifgf_fno_error(const ifgf_fno* fno); /* NULL: the thread's last error */
/* FP32 master weights, flat (ifgf_fno_num_params() values) */
int ifgf_fno_get_params(const ifgf_fno* fno, float* out);
int ifgf_fno_set_params(ifgf_fno* fno, const float* in);
/* Adds a training sample: right-hand side v and the solved density (host,
 * N complex each, patch order) of a plan's curve; its features are formed on
 * the device and kept with the network. */
int ifgf_fno_add_sample(ifgf_fno* fno, ifgf_plan* plan, const double* v, const double* density);
int ifgf_fno_num_samples(const ifgf_fno* fno);
int ifgf_fno_clear_samples(ifgf_fno* fno);
/* Sample s's features X [N][11], prior multiplier [N], node parameters t [N]
 * and parameter quadrature weights [N] (host; any may be NULL). */
int ifgf_fno_sample_data(const ifgf_fno* fno, int s, long long* N, float* X, float* mphys, double* t, double* wt);
/* Mean over samples ids[0..n) of ||M(v) - density||^2 / ||density||^2, and
 * (grad != NULL) its gradient with respect to the master weights. */
int ifgf_fno_loss_grad(ifgf_fno* fno, const int* ids, int n, double* loss, float* grad);
/* Adam (lr, 0.9, 0.999, 1e-8) over samples ids[0..n) for `epochs` seeded
 * shuffles in minibatches of `batch`; loss_hist[e] = epoch e's mean loss. */
int ifgf_fno_train(ifgf_fno* fno, const int* ids, int n, int epochs, int batch, double lr, unsigned long long seed,
                   double* loss_hist);
/* z = M(v) on a plan's curve; flags IFGF_PTR_DEVICE (device pointers,
 * asynchronous on stream) or host pointers (synchronous). */
int ifgf_fno_apply(ifgf_fno* fno, ifgf_plan* plan, const double* v, double* z, int flags, void* stream);
/* BF16 MFMA GEMM of the network (C = A B, FP32 row-major host arrays; A
 * given as K x M when trans_a, B as N x K when trans_b), for tests. */
int ifgf_fno_gemm_check(int M, int N, int K, int trans_a, int trans_b, const float* A, const float* B, float* C);
/* GMRES with the learned network (the plan's default operator; host
 * pointers; x holds the initial guess): mode IFGF_FNO_GUESS replaces a zero
 * x by M(b); IFGF_FNO_PRECONDITION runs flexible GMRES(restart) with
 * z_j = M(v_j), and on stagnation (an estimate above ratio x the one `window`
 * iterations earlier) falls back to plain GMRES (the identity preconditioner)
 * from the current iterate (*fallback = 1). Every estimate is |w(k)| / ||b||
 * (the reference divides by ||b - A x0||, opGMRES.hpp:105,207, which would
 * discard what the initial density gains); stops below tol or after maxit
 * Arnoldi steps in all. relres: the true FP64 ||b - A x|| / ||b||. */
int ifgf_fno_gmres(ifgf_plan* plan, ifgf_fno* fno, const double* rhs, double* x, int restart, double tol, int maxit,
                   int mode, int window, double ratio, int* iters, int* fallback, double* relres, double* hist,
                   int hist_cap);

/* Generic host GMRES on any operator: the reference's Householder GMRES
 * (opGMRES.hpp:68-269) with the mapping of LinearOperator.hpp:54 as a C
 * callback. Returns IFGF_OK / IFGF_ERR_NOT_CONVERGED; err is the estimate. */
typedef void (*ifgf_apply_fn)(void* ctx, const double* x, double* y);
int ifgf_gmres_host(ifgf_apply_fn apply, void* ctx, long long n, const double* rhs, double* x,
                    int restart, double tol, int maxit, int* iters, double* err);

/* H0^(1)(x) and H1^(1)(x) on the GPU for x[0..n) (replaces
 * sp_bessel::hankelH1(0|1, x), besselFunctions.h:312-341). Host pointers;
 * h0 / h1 receive n interleaved complex values. */
int ifgf_hankel01(const double* x, long long n, double* h0, double* h1);

/* The discretised curve's parametrisation on the host: replaces Curve::xt,
 * yt, xpt, ypt (Geometry/Curve.hpp:97-146, Curve.cpp:105-138; the star has no
 * reference class) at t[0..n); any output may be NULL. curve_type / params as
 * in ifgf_surface. */
int ifgf_curve_eval(int curve_type, const double* params, int nparams, const double* t, long long n, double* x,
                    double* y, double* xp, double* yp);

/* (x(a) - x(a - d), y(a) - y(a - d)) at n pairs, formed with trigonometric
 * difference identities (accurate relative to the difference itself); the
 * reference subtracts Cartesian coordinates in the RP precomputation
 * (ForwardMap.tpp:355-365), which this replaces. */
int ifgf_curve_diff(int curve_type, const double* params, int nparams, const double* a, const double* d,
                    long long n, double* dx, double* dy);

/* Per-stage GPU timing of applications: hipEvents recorded on the stream each
 * application's kernels run on, without host synchronisation (the hipEvent
 * timers that replace the reference's std::chrono prints,
 * ForwardMap.tpp:545-563). enable != 0 clears and starts recording;
 * ifgf_plan_stage_times waits for the recorded applications and returns the
 * accumulated time of [0] intensities+sort, [1] near field, [2] leaf cones,
 * [3] propagation, [4] interpolation, [5] correction+unsort, and in *count
 * the number of applications recorded. Each stage time is the span of its
 * kernels on the application's stream. Recording the events costs time of
 * its own (several microseconds per stage), so applications timed for
 * throughput should run with profiling off. */
int ifgf_plan_profile(ifgf_plan* plan, int enable);
int ifgf_plan_stage_times(ifgf_plan* plan, double times_ms[6], long long* count);

/* Wall-clock seconds of plan creation by step: [0] discretisation and
 * near-singular detection, [1] singular / near-singular correction, [2] Morton
 * sort and tree tables, [3] near-field exclusions and cone records, [4] device
 * tables, [5] stored operators (kernel evaluations and the level-push
 * tables), [6] autotuning, [7] total. No reference counterpart (the reference
 * times nothing inside precomps_and_setup). */
int ifgf_plan_setup_times(const ifgf_plan* plan, double seconds[8]);

#ifdef __cplusplus
}
#endif
#endif /* IFGF2D_H */
