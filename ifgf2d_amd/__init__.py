"""MI355X IFGF forward map for 2D Helmholtz boundary integral equations.

Python host mirror of manuelarturosantana/IFGF2D's operator interface, calling
the C ABI of libifgf2d.so (include/ifgf2d.h):

* ``ForwardMap``      -- ForwardMap<Np, Formulation, Ps, Pang> (ForwardMap/ForwardMap.hpp:49-177)
                          with precomps_and_setup folded into construction;
                          ``compute_Ax_acc`` is the GPU forward map.
* ``LinearOperator``  -- ForwardMap/LinearOperator.hpp:27-60: rows/cols plus a
                          mapping(x, out) callable, applied by ``op * x``.
* ``gmres``           -- the reference's Householder GMRES (opGMRES.hpp:68-269)
                          on any LinearOperator (host), and ``ForwardMap.solve``
                          for the device-resident GMRES.

There is no CPU fallback: importing works without a GPU (plans can be built
host-only for inspection), but every compute call runs the HIP kernels and
raises if the library or the device is missing.
"""
import ctypes
import math
import os

import numpy as np

__all__ = [
    "ForwardMap", "LinearOperator", "ObstacleGroup", "LearnedPreconditioner", "fno_gemm_check", "gmres", "hankel01", "morton_sort", "IFGFError",
    "circle", "kite", "star", "equal_parameter_lims", "SINGLE_LAYER", "DOUBLE_LAYER", "COMBINED",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libifgf2d.so")

SINGLE_LAYER, DOUBLE_LAYER, COMBINED = 0, 1, 2
FP64, MIXED = 0, 1
_FORMS = {"single": SINGLE_LAYER, "double": DOUBLE_LAYER, "combined": COMBINED,
          SINGLE_LAYER: SINGLE_LAYER, DOUBLE_LAYER: DOUBLE_LAYER, COMBINED: COMBINED}
_PREC = {"fp64": FP64, "mixed": MIXED, FP64: FP64, MIXED: MIXED}

IFGF_OK, IFGF_ERR_NOT_CONVERGED = 0, 6
IFGF_PLAN_HOST_ONLY, IFGF_PLAN_ON_THE_FLY, IFGF_PLAN_AUTOTUNE, IFGF_PTR_DEVICE = 1, 2, 4, 1
IFGF_APPLY_FP64 = 2
IFGF_PLAN_DEVICE_SORT = 8
IFGF_PLAN_OVERLAP = 16
IFGF_PLAN_DEVICE_PRECOMP = 32
IFGF_PLAN_DEVICE_GEOMETRY = 64
IFGF_PLAN_DEVICE_TREE = 128
IFGF_PLAN_NO_MFMA = 256
IFGF_FNO_GUESS, IFGF_FNO_PRECONDITION, IFGF_FNO_FIXED_MULTIPLIER = 1, 2, 4


class IFGFError(RuntimeError):
    pass


class _Surface(ctypes.Structure):
    _fields_ = [("curve_type", ctypes.c_int), ("curve_params", ctypes.POINTER(ctypes.c_double)),
                ("ncurve_params", ctypes.c_int), ("patch_lims", ctypes.POINTER(ctypes.c_double)),
                ("npatches", ctypes.c_int), ("Np", ctypes.c_int), ("delta", ctypes.c_double),
                ("rp_p", ctypes.c_double)]


class _Params(ctypes.Structure):
    _fields_ = [("formulation", ctypes.c_int), ("precision", ctypes.c_int), ("k", ctypes.c_double),
                ("eta", ctypes.c_double), ("Ps", ctypes.c_int), ("Pang", ctypes.c_int),
                ("nlevels", ctypes.c_int), ("nrhs_max", ctypes.c_int), ("flags", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_longlong)
_ip = ctypes.POINTER(ctypes.c_int)
APPLY_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _dp, _dp)

EXPORTS = [
    "ifgf_plan_create", "ifgf_plan_destroy", "ifgf_last_error", "ifgf_plan_error", "ifgf_plan_size",
    "ifgf_plan_points", "ifgf_plan_stats", "ifgf_plan_sorting", "ifgf_plan_level_cones",
    "ifgf_plan_correction", "ifgf_apply", "ifgf_apply_direct", "ifgf_propagate", "ifgf_gmres",
    "ifgf_fgmres", "ifgf_gmres_host", "ifgf_hankel01", "ifgf_plan_profile", "ifgf_plan_stage_times",
    "ifgf_morton_sort", "ifgf_apply_graph", "ifgf_group_create", "ifgf_group_launch",
    "ifgf_group_synchronize", "ifgf_group_destroy", "ifgf_gmres_multi", "ifgf_field_create",
    "ifgf_field_apply", "ifgf_field_stats", "ifgf_field_destroy", "ifgf_plan_setup_times",
    "ifgf_refine_multi", "ifgf_curve_eval", "ifgf_curve_diff", "ifgf_plan_solver_times",
    "ifgf_propagate_noalloc", "ifgf_fno_num_params", "ifgf_fno_create", "ifgf_fno_destroy", "ifgf_fno_error",
    "ifgf_fno_get_params", "ifgf_fno_set_params", "ifgf_fno_add_sample", "ifgf_fno_num_samples",
    "ifgf_fno_clear_samples", "ifgf_fno_sample_data", "ifgf_fno_loss_grad", "ifgf_fno_train", "ifgf_fno_apply",
    "ifgf_fno_gemm_check", "ifgf_fno_gmres", "ifgf_group_stats", "ifgf_autotune_cone_order",
]

_LIB = None


def lib():
    """Load libifgf2d.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise IFGFError(f"{LIB_PATH} is missing: build it with `make -C ifgf2d_amd` "
                            "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.ifgf_plan_create.argtypes = [ctypes.POINTER(_Surface), ctypes.POINTER(_Params),
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.ifgf_plan_destroy.argtypes = [ctypes.c_void_p]
        L.ifgf_last_error.restype = ctypes.c_char_p
        L.ifgf_plan_error.argtypes = [ctypes.c_void_p]
        L.ifgf_plan_error.restype = ctypes.c_char_p
        L.ifgf_plan_size.argtypes = [ctypes.c_void_p]
        L.ifgf_plan_size.restype = ctypes.c_longlong
        L.ifgf_plan_points.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.ifgf_plan_stats.argtypes = [ctypes.c_void_p, _lp]
        L.ifgf_plan_sorting.argtypes = [ctypes.c_void_p, _lp]
        L.ifgf_plan_level_cones.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _lp]
        L.ifgf_plan_level_cones.restype = ctypes.c_longlong
        L.ifgf_plan_correction.argtypes = [ctypes.c_void_p, _lp, _ip, _dp]
        L.ifgf_plan_correction.restype = ctypes.c_longlong
        L.ifgf_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p]
        L.ifgf_apply_graph.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p]
        L.ifgf_group_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.ifgf_group_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ifgf_group_synchronize.argtypes = [ctypes.c_void_p]
        L.ifgf_group_destroy.argtypes = [ctypes.c_void_p]
        L.ifgf_apply_direct.argtypes = [ctypes.c_void_p, _dp, _lp, ctypes.c_longlong, _dp]
        L.ifgf_propagate.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_longlong, _dp, _dp]
        L.ifgf_gmres.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, _ip, _dp, _dp, ctypes.c_int]
        L.ifgf_fgmres.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip, _ip, _dp, _dp,
                                  ctypes.c_int]
        L.ifgf_gmres_multi.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, _ip, _dp]
        L.ifgf_refine_multi.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip, _ip, _dp]
        L.ifgf_gmres_host.argtypes = [APPLY_FN, ctypes.c_void_p, ctypes.c_longlong, _dp, _dp,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, _dp]
        L.ifgf_hankel01.argtypes = [_dp, ctypes.c_longlong, _dp, _dp]
        L.ifgf_morton_sort.argtypes = [_dp, _dp, ctypes.c_longlong, ctypes.c_int, _lp, _lp]
        L.ifgf_curve_eval.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_longlong, _dp, _dp, _dp,
                                      _dp]
        L.ifgf_curve_diff.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, ctypes.c_longlong, _dp, _dp]
        L.ifgf_plan_solver_times.argtypes = [ctypes.c_void_p, _dp]
        L.ifgf_propagate_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _fp = ctypes.POINTER(ctypes.c_float)
        L.ifgf_fno_num_params.restype = ctypes.c_longlong
        L.ifgf_fno_create.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_void_p)]
        L.ifgf_fno_destroy.argtypes = [ctypes.c_void_p]
        L.ifgf_fno_error.argtypes = [ctypes.c_void_p]
        L.ifgf_fno_error.restype = ctypes.c_char_p
        L.ifgf_fno_get_params.argtypes = [ctypes.c_void_p, _fp]
        L.ifgf_fno_set_params.argtypes = [ctypes.c_void_p, _fp]
        L.ifgf_fno_add_sample.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp, _dp]
        L.ifgf_fno_num_samples.argtypes = [ctypes.c_void_p]
        L.ifgf_fno_clear_samples.argtypes = [ctypes.c_void_p]
        L.ifgf_fno_sample_data.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _fp, _fp, _dp, _dp]
        L.ifgf_fno_loss_grad.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, _dp, _fp]
        L.ifgf_fno_train.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_ulonglong, _dp]
        L.ifgf_fno_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p]
        L.ifgf_fno_gemm_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp,
                                          _fp]
        L.ifgf_fno_gmres.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip, _ip, _dp, _dp,
                                     ctypes.c_int]
        L.ifgf_group_stats.argtypes = [ctypes.c_void_p, _lp]
        L.ifgf_autotune_cone_order.argtypes = [ctypes.POINTER(_Surface), ctypes.POINTER(_Params), ctypes.c_double, _ip,
                                               _ip, _dp]
        L.ifgf_plan_profile.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ifgf_plan_stage_times.argtypes = [ctypes.c_void_p, _dp, _lp]
        L.ifgf_plan_setup_times.argtypes = [ctypes.c_void_p, _dp]
        L.ifgf_field_create.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.ifgf_field_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_void_p]
        L.ifgf_field_stats.argtypes = [ctypes.c_void_p, _lp]
        L.ifgf_field_destroy.argtypes = [ctypes.c_void_p]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


# ------------------------------------------------------------------ curves ---
def circle(cx=0.0, cy=0.0, R=1.0):
    """Geometry/Curve.hpp:106-125."""
    return (0, [cx, cy, R])


def kite(A=1.0, B=0.65, C=-0.65, D=1.5):
    """Geometry/Curve.hpp:128-146 (Colton-Kress kite)."""
    return (1, [A, B, C, D])


def star(cx=0.0, cy=0.0, R=1.0, a=(), b=()):
    """r(t) = R (1 + sum_m a[m] cos(m t) + b[m] sin(m t)), m = 0.. (no reference class)."""
    n = max(len(a), len(b))
    a = list(a) + [0.0] * (n - len(a))
    b = list(b) + [0.0] * (n - len(b))
    return (2, [cx, cy, R, float(n)] + a + b)


def equal_parameter_lims(npatches):
    return np.linspace(0.0, 2.0 * np.pi, npatches + 1)


def _curve_eval(curve, t):
    typ, p = curve
    if typ == 0:
        return p[0] + p[2] * np.cos(t), p[1] + p[2] * np.sin(t)
    if typ == 1:
        return p[0] * np.cos(t) + p[1] * np.cos(2 * t) + p[2], p[3] * np.sin(t)
    if typ != 2 or len(p) < 4:
        raise IFGFError("unknown curve type or too few curve parameters")
    n = int(p[3])
    r = np.ones_like(t)
    for m in range(n):
        r = r + p[4 + m] * np.cos(m * t) + p[4 + n + m] * np.sin(m * t)
    return p[0] + p[2] * r * np.cos(t), p[1] + p[2] * r * np.sin(t)


def curve_eval(curve, t):
    """(x, y, x', y') of a curve at the parameters t (Curve::xt/yt/xpt/ypt, host)."""
    typ, p = curve
    prm = np.ascontiguousarray(p, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = [np.empty_like(t) for _ in range(4)]
    rc = lib().ifgf_curve_eval(typ, _d(prm), prm.size, _d(t), t.size, *[_d(o) for o in out])
    if rc != IFGF_OK:
        raise IFGFError(lib().ifgf_last_error().decode())
    return tuple(out)


def curve_diff(curve, a, d):
    """(x(a) - x(a - d), y(a) - y(a - d)) from difference identities (host)."""
    typ, p = curve
    prm = np.ascontiguousarray(p, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    dx, dy = np.empty_like(a), np.empty_like(a)
    rc = lib().ifgf_curve_diff(typ, _d(prm), prm.size, _d(a), _d(d), a.size, _d(dx), _d(dy))
    if rc != IFGF_OK:
        raise IFGFError(lib().ifgf_last_error().decode())
    return dx, dy


def default_delta(curve, lims, frac=0.75):
    """Near-singular distance: frac x the shortest patch chord."""
    x, y = _curve_eval(curve, np.asarray(lims, dtype=np.float64))
    return frac * float(np.min(np.hypot(np.diff(x), np.diff(y))))


# -------------------------------------------------------------- operators ---
class ForwardMap:
    """The IFGF-accelerated boundary-integral operator on one closed curve.

    Mirrors ForwardMap<Np, Formulation, Ps, Pang>(delta, curves, ..., p) followed
    by precomps_and_setup(k, nlevels) (ForwardMap.tpp:15-35, 542-565).
    formulation: "single" (the reference's SingleLayer S), "double" (1/2 I + K)
    or "combined" (1/2 I + K - i eta S, the sound-soft CFIE).
    on_the_fly: re-evaluate the near-field and leaf kernels in every application
    instead of storing them once at plan creation (less memory, slower).
    autotune: time the candidate lane counts of the stored-path kernels once.
    precision: "fp64", or "mixed" (FP32 leaf operator, cone coefficients, propagation and
    interpolation; the FP64 operator stays available through compute_Ax_acc(x, fp64=True)).
    device_sort: Morton box assignment and point sort on the GPU (checked against the
    host's stable sort).
    overlap: run the stored near field (and the identity / correction half of the final
    sum) on a second stream beside the level pushes (the autotuner also tries it).
    device_precomp: build the rectangular-polar singular / near-singular correction on the
    GPU instead of the host.
    device_geometry: run the near-singular search of the discretisation on the GPU.
    device_tree: build the tree's relevant cone segments (interpolation entries, protoboxes,
        numbering) on the GPU; the tables equal the host's.
    device_setup: all four device setup steps (sort, correction, near-singular search, cones).
    mfma: blocks of 8 columns (nrhs_max >= 8) run the near field and leaf operator on the
        matrix cores (False: the VALU batched kernels).
    """

    def __init__(self, curve, patch_lims, k, nlevels, Np=10, formulation="combined", eta=None,
                 delta=None, p=4.0, Ps=5, Pang=5, precision="fp64", nrhs_max=1, host_only=False,
                 on_the_fly=False, autotune=False, device_sort=False, overlap=False, device_precomp=False,
                 device_geometry=False, device_setup=False, device_tree=False, mfma=True):
        L = lib()
        self.curve = curve
        self.lims = np.ascontiguousarray(patch_lims, dtype=np.float64)
        self.k = float(k)
        self.eta = float(k if eta is None else eta)
        self.formulation = _FORMS[formulation]
        self.delta = default_delta(curve, self.lims) if delta is None else float(delta)
        self._cp = np.ascontiguousarray(curve[1], dtype=np.float64)
        surf = _Surface(curve[0], _d(self._cp), self._cp.size, _d(self.lims), self.lims.size - 1,
                        Np, self.delta, float(p))
        flags = ((IFGF_PLAN_HOST_ONLY if host_only else 0) | (IFGF_PLAN_ON_THE_FLY if on_the_fly else 0)
                 | (IFGF_PLAN_AUTOTUNE if autotune else 0) | (IFGF_PLAN_DEVICE_SORT if device_sort else 0)
                 | (IFGF_PLAN_OVERLAP if overlap else 0) | (IFGF_PLAN_DEVICE_PRECOMP if device_precomp else 0)
                 | (IFGF_PLAN_DEVICE_GEOMETRY if device_geometry else 0)
                 | (IFGF_PLAN_DEVICE_TREE if device_tree else 0) | (0 if mfma else IFGF_PLAN_NO_MFMA))
        if device_setup:
            flags |= (IFGF_PLAN_DEVICE_SORT | IFGF_PLAN_DEVICE_PRECOMP | IFGF_PLAN_DEVICE_GEOMETRY
                      | IFGF_PLAN_DEVICE_TREE)
        prm = _Params(self.formulation, _PREC[precision], self.k, self.eta, Ps, Pang, nlevels,
                      nrhs_max, flags)
        h = ctypes.c_void_p()
        rc = L.ifgf_plan_create(ctypes.byref(surf), ctypes.byref(prm), ctypes.byref(h))
        if rc != IFGF_OK:
            raise IFGFError(f"ifgf_plan_create failed ({rc}): {L.ifgf_last_error().decode()}")
        self._h = h
        self.N = L.ifgf_plan_size(h)
        self.Np, self.Ps, self.Pang, self.nlevels = Np, Ps, Pang, nlevels
        self.npatches = self.lims.size - 1

    def close(self):
        if getattr(self, "_h", None):
            lib().ifgf_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc not in (IFGF_OK,):
            raise IFGFError(f"{what} failed ({rc}): {lib().ifgf_plan_error(self._h).decode()}")

    # ----- introspection
    def points(self):
        x, y, nx, ny, w = (np.empty(self.N) for _ in range(5))
        self._check(lib().ifgf_plan_points(self._h, _d(x), _d(y), _d(nx), _d(ny), _d(w)), "points")
        return x, y, nx, ny, w

    def _stats(self):
        s = np.empty(20, dtype=np.int64)
        self._check(lib().ifgf_plan_stats(self._h, s.ctypes.data_as(_lp)), "stats")
        return s.tolist()

    def stats(self):
        """Work counts of one application (near pairs, leaf evaluations, ...)."""
        return dict(zip(("near", "leaf", "prop", "interp", "cones", "corr", "nlevels", "N"),
                        self._stats()[:8]))

    def extra_stats(self):
        s = self._stats()
        return {"leaf_cones": s[8], "prop_cones": s[4] - s[8], "stored": bool(s[9]),
                "near_stored": s[10], "stored_bytes": s[11],
                "lanes": {"near": s[12], "leaf": s[13], "push_task_entries": s[14], "finalize": s[15],
                          "near_on_second_stream": bool(s[16]), "near_fork": s[16]},
                "mixed": bool(s[17]), "ptab_entries": s[18], "ifac_entries": s[19]}

    def sorting(self):
        s = np.empty(self.N, dtype=np.int64)
        self._check(lib().ifgf_plan_sorting(self._h, s.ctypes.data_as(_lp)), "sorting")
        return s

    def level_cones(self, level):
        n = lib().ifgf_plan_level_cones(self._h, level, None, None)
        if n < 0:
            raise IFGFError(lib().ifgf_last_error().decode())
        b, q = np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int64)
        lib().ifgf_plan_level_cones(self._h, level, b.ctypes.data_as(_lp), q.ctypes.data_as(_lp))
        return b, q

    def correction(self):
        n = lib().ifgf_plan_correction(self._h, None, None, None)
        ptr = np.empty(self.N + 1, dtype=np.int64)
        patch = np.empty(n, dtype=np.int32)
        w = np.empty(2 * n * self.Np)
        lib().ifgf_plan_correction(self._h, ptr.ctypes.data_as(_lp), patch.ctypes.data_as(_ip), _d(w))
        return ptr, patch, w.view(np.complex128).reshape(n, self.Np)

    # ----- the forward map
    def compute_Ax_acc(self, x, fp64=False):
        """y = A x on the GPU for host vectors; x may be (N,) or (N, nrhs) (column major).
        fp64: on a mixed-precision plan, apply the all-FP64 operator."""
        x = np.asarray(x, dtype=np.complex128)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        xf = np.asfortranarray(x.reshape(self.N, nrhs))
        y = np.empty_like(xf, order="F")
        self._check(lib().ifgf_apply(self._h, xf.ctypes.data, y.ctypes.data, nrhs, self.N,
                                     IFGF_APPLY_FP64 if fp64 else 0, None), "ifgf_apply")
        return y.reshape(x.shape)

    __call__ = compute_Ax_acc

    def apply_device(self, x_ptr, y_ptr, nrhs=1, ld=None, stream=None, fp64=False):
        """Asynchronous apply on device pointers (e.g. torch tensors' data_ptr())."""
        flags = IFGF_PTR_DEVICE | (IFGF_APPLY_FP64 if fp64 else 0)
        self._check(lib().ifgf_apply(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), nrhs,
                                     self.N if ld is None else ld, flags,
                                     ctypes.c_void_p(stream) if stream else None), "ifgf_apply")

    def apply_graph(self, x_ptr, y_ptr, nrhs=1, ld=None, stream=None, fp64=False):
        """apply_device replayed as one captured HIP graph (captured on the first call
        with these pointers and shape)."""
        self._check(lib().ifgf_apply_graph(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), nrhs,
                                           self.N if ld is None else ld, IFGF_APPLY_FP64 if fp64 else 0,
                                           ctypes.c_void_p(stream) if stream else None), "ifgf_apply_graph")

    def compute_Ax_unacc(self, x, targets):
        """Unaccelerated operator (direct sum + correction) at a target subset."""
        x = _c(x)
        t = np.ascontiguousarray(targets, dtype=np.int64)
        out = np.empty(t.size, dtype=np.complex128)
        self._check(lib().ifgf_apply_direct(self._h, _d(x.view(np.float64)), t.ctypes.data_as(_lp),
                                            t.size, _d(out.view(np.float64))), "ifgf_apply_direct")
        return out

    def propagate_unacc(self, tx, ty, density):
        """Field of the layer potential at off-surface points (ForwardMap.tpp:702-727)."""
        tx = np.ascontiguousarray(tx, dtype=np.float64)
        ty = np.ascontiguousarray(ty, dtype=np.float64)
        d = _c(density)
        out = np.empty(tx.size, dtype=np.complex128)
        self._check(lib().ifgf_propagate(self._h, _d(tx), _d(ty), tx.size, _d(d.view(np.float64)),
                                         _d(out.view(np.float64))), "ifgf_propagate")
        return out

    def propagate_noalloc_device(self, tx_ptr, ty_ptr, nt, density_ptr, out_ptr, stream=None):
        """propagate_unacc on device pointers through the library's fixed scratch
        (ifgf_propagate_noalloc: allocates nothing; asynchronous)."""
        self._check(lib().ifgf_propagate_noalloc(self._h, ctypes.c_void_p(tx_ptr), ctypes.c_void_p(ty_ptr), nt,
                                                 ctypes.c_void_p(density_ptr), ctypes.c_void_p(out_ptr),
                                                 ctypes.c_void_p(stream) if stream else None),
                    "ifgf_propagate_noalloc")

    def field(self, tx, ty, nlevels=None, Ps=5, Pang=5, precision="fp64"):
        """IFGF-accelerated field evaluator at the points (tx, ty): the sum of
        propagate_unacc (ForwardMap.tpp:702-727) through an IFGF tree of its own.
        nlevels defaults to this plan's levels plus the number of doublings from the
        points' bounding box to that of the points and targets (the leaf size stays
        about the same)."""
        return Field(self, tx, ty, nlevels, Ps, Pang, precision)

    def solve(self, rhs, restart=50, tol=1e-8, maxit=1000, x0=None):
        """Device GMRES (ForwardMap::solve, ForwardMap.tpp:658-697).

        Returns (x, iterations, true relative residual, residual-estimate history)."""
        b = _c(rhs)
        x = np.zeros_like(b) if x0 is None else _c(x0).copy()
        it, rr = ctypes.c_int(), ctypes.c_double()
        hist = np.zeros(maxit)
        rc = lib().ifgf_gmres(self._h, _d(b.view(np.float64)), _d(x.view(np.float64)), restart, tol,
                              maxit, ctypes.byref(it), ctypes.byref(rr), _d(hist), hist.size)
        if rc not in (IFGF_OK, IFGF_ERR_NOT_CONVERGED):
            self._check(rc, "ifgf_gmres")
        return x, it.value, rr.value, hist[: it.value]

    def solve_multi(self, rhs, restart=50, tol=1e-8, maxit=1000, x0=None):
        """GMRES for the columns of rhs (N x nrhs) at once, lockstep with batched
        applications (ifgf_gmres_multi). Returns (X, iterations per column, true
        relative residual per column)."""
        B = np.asfortranarray(np.asarray(rhs, dtype=np.complex128).reshape(self.N, -1))
        nr = B.shape[1]
        X = np.zeros_like(B, order="F") if x0 is None else np.asfortranarray(np.array(x0, dtype=np.complex128))
        its = np.zeros(nr, dtype=np.int32)
        rr = np.zeros(nr)
        rc = lib().ifgf_gmres_multi(self._h, B.ctypes.data_as(_dp), X.ctypes.data_as(_dp), nr, restart, tol, maxit,
                                    its.ctypes.data_as(_ip), _d(rr))
        if rc not in (IFGF_OK, IFGF_ERR_NOT_CONVERGED):
            self._check(rc, "ifgf_gmres_multi")
        return X, its, rr

    def solve_multi_refined(self, rhs, restart=50, tol=1e-8, maxit=1000, max_outer=10, inner_floor=1e-5,
                            x0=None):
        """Mixed-precision refinement for the columns of rhs (ifgf_refine_multi):
        lockstep GMRES corrections with the plan's default operator (mixed precision
        on a mixed plan), FP64 residuals. Returns (X, refinement steps, inner
        iterations per column, true FP64 relative residual per column)."""
        B = np.asfortranarray(np.asarray(rhs, dtype=np.complex128).reshape(self.N, -1))
        nr = B.shape[1]
        X = np.zeros_like(B, order="F") if x0 is None else np.asfortranarray(np.array(x0, dtype=np.complex128))
        its = np.zeros(nr, dtype=np.int32)
        rr = np.zeros(nr)
        outer = ctypes.c_int(0)
        rc = lib().ifgf_refine_multi(self._h, B.ctypes.data_as(_dp), X.ctypes.data_as(_dp), nr, restart, tol,
                                     maxit, max_outer, inner_floor, ctypes.byref(outer), its.ctypes.data_as(_ip),
                                     _d(rr))
        if rc not in (IFGF_OK, IFGF_ERR_NOT_CONVERGED):
            self._check(rc, "ifgf_refine_multi")
        return X, outer.value, its, rr

    def solve_fgmres(self, rhs, restart=50, tol=1e-8, maxit=200, inner_restart=50, inner_floor=1e-5,
                     x0=None):
        """Flexible GMRES with FP64 refinement: FP64 outer Arnoldi with the all-FP64
        operator, inner GMRES cycles with the plan's default operator (mixed precision
        on a mixed plan). Returns (x, outer iterations, inner iterations, true relative
        residual, outer residual-estimate history)."""
        b = _c(rhs)
        x = np.zeros_like(b) if x0 is None else _c(x0).copy()
        it, iit, rr = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        hist = np.zeros(maxit)
        rc = lib().ifgf_fgmres(self._h, _d(b.view(np.float64)), _d(x.view(np.float64)), restart, tol,
                               maxit, inner_restart, inner_floor, ctypes.byref(it), ctypes.byref(iit),
                               ctypes.byref(rr), _d(hist), hist.size)
        if rc not in (IFGF_OK, IFGF_ERR_NOT_CONVERGED):
            self._check(rc, "ifgf_fgmres")
        return x, it.value, iit.value, rr.value, hist[: it.value]

    def profile(self, enable=True):
        self._check(lib().ifgf_plan_profile(self._h, 1 if enable else 0), "profile")

    def stage_times(self):
        t = np.zeros(6)
        c = ctypes.c_longlong()
        self._check(lib().ifgf_plan_stage_times(self._h, _d(t), ctypes.byref(c)), "stage_times")
        return dict(zip(("gather", "near", "leaf", "propagation", "interpolation", "finalize"),
                        t.tolist())), c.value

    def setup_times(self):
        """Seconds spent in each plan-creation step."""
        t = np.zeros(8)
        self._check(lib().ifgf_plan_setup_times(self._h, _d(t)), "setup_times")
        return dict(zip(("geometry", "correction", "tree", "exclusions", "device_tables", "stored_operators",
                         "autotune", "total"), t.tolist()))

    def solver_times(self):
        """Seconds of the last solve/solve_fgmres call by phase."""
        t = np.zeros(4)
        self._check(lib().ifgf_plan_solver_times(self._h, _d(t)), "solver_times")
        return dict(zip(("workspace", "host_waits", "inner_cycles", "total"), t.tolist()))

    def solve_learned(self, rhs, fno, restart=50, tol=1e-8, maxit=1000, guess=True, precondition=False,
                      window=8, ratio=0.5, x0=None, fixed_multiplier=False):
        """GMRES with the learned network (ifgf_fno_gmres): guess -> initial density
        M(b); precondition -> flexible right preconditioner z = M(v), falling back to
        plain GMRES when the estimate stops falling (ratio over window iterations).
        Returns (x, Arnoldi steps, fell back, true FP64 relative residual, estimates)."""
        b = _c(rhs)
        x = np.zeros_like(b) if x0 is None else _c(x0).copy()
        it, fb, rr = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        hist = np.zeros(maxit)
        mode = ((IFGF_FNO_GUESS if guess else 0) | (IFGF_FNO_PRECONDITION if precondition else 0)
                | (IFGF_FNO_FIXED_MULTIPLIER if fixed_multiplier else 0))
        rc = lib().ifgf_fno_gmres(self._h, fno._h, _d(b.view(np.float64)), _d(x.view(np.float64)), restart, tol,
                                  maxit, mode, window, ratio, ctypes.byref(it), ctypes.byref(fb), ctypes.byref(rr),
                                  _d(hist), hist.size)
        if rc not in (IFGF_OK, IFGF_ERR_NOT_CONVERGED):
            self._check(rc, "ifgf_fno_gmres")
        return x, it.value, bool(fb.value), rr.value, hist[: it.value]

    def linear_operator(self):
        """The LinearOperator that ForwardMap::solve builds (ForwardMap.tpp:673-682)."""
        def mapping(x, out):
            out[:] = self.compute_Ax_acc(x)
        return LinearOperator(self.N, self.N, mapping)


class Field:
    """IFGF-accelerated evaluation of a ForwardMap's layer potential at separate
    target points (ifgf_field_*). evaluate(density) returns, per target,
    sum_s G(target, x_s) density_s w_s over the plan's points (coincident points
    contribute zero); G, k and eta are the plan's."""

    def __init__(self, fm, tx, ty, nlevels=None, Ps=5, Pang=5, precision="fp64"):
        self.tx = np.ascontiguousarray(tx, dtype=np.float64).ravel()
        self.ty = np.ascontiguousarray(ty, dtype=np.float64).ravel()
        if self.tx.size != self.ty.size:
            raise IFGFError("tx and ty differ in length")
        self.M, self.N = self.tx.size, fm.N
        if nlevels is None:
            x, y = fm.points()[:2]
            s0 = max(np.ptp(x), np.ptp(y))
            s1 = max(np.ptp(np.concatenate([x, self.tx])), np.ptp(np.concatenate([y, self.ty])))
            nlevels = fm.nlevels + max(0, int(np.round(np.log2(s1 / s0))))
        self.nlevels = int(nlevels)
        h = ctypes.c_void_p()
        rc = lib().ifgf_field_create(fm._h, _d(self.tx), _d(self.ty), self.M, self.nlevels, Ps, Pang,
                                     _PREC[precision], ctypes.byref(h))
        if rc != IFGF_OK:
            raise IFGFError(f"ifgf_field_create failed ({rc}): {lib().ifgf_last_error().decode()}")
        self._h = h

    def evaluate(self, density, fp64=False):
        d = _c(density)
        if d.size != self.N:
            raise IFGFError("density length differs from the plan's point count")
        out = np.empty(self.M, dtype=np.complex128)
        rc = lib().ifgf_field_apply(self._h, d.ctypes.data, out.ctypes.data, IFGF_APPLY_FP64 if fp64 else 0, None)
        if rc != IFGF_OK:
            raise IFGFError(f"ifgf_field_apply failed ({rc}): {lib().ifgf_last_error().decode()}")
        return out

    __call__ = evaluate

    def apply_device(self, d_ptr, out_ptr, stream=None, fp64=False):
        """Asynchronous evaluation on device pointers (density N, out M complex128)."""
        flags = IFGF_PTR_DEVICE | (IFGF_APPLY_FP64 if fp64 else 0)
        rc = lib().ifgf_field_apply(self._h, ctypes.c_void_p(d_ptr), ctypes.c_void_p(out_ptr), flags,
                                    ctypes.c_void_p(stream) if stream else None)
        if rc != IFGF_OK:
            raise IFGFError(f"ifgf_field_apply failed ({rc}): {lib().ifgf_last_error().decode()}")

    def stats(self):
        s = np.empty(8, dtype=np.int64)
        rc = lib().ifgf_field_stats(self._h, s.ctypes.data_as(_lp))
        if rc != IFGF_OK:
            raise IFGFError(lib().ifgf_last_error().decode())
        return dict(zip(("targets", "near", "leaf", "prop", "interp", "cones", "nlevels", "stored_bytes"),
                        s.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            lib().ifgf_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ObstacleGroup:
    """Several obstacles' forward maps as one captured launch (ifgf_group_*):
    maps[i] applied to the device vector x_ptrs[i] into y_ptrs[i]."""

    def __init__(self, maps, x_ptrs, y_ptrs):
        n = len(maps)
        self.maps = list(maps)  # the plans must outlive the group
        P = (ctypes.c_void_p * n)(*[m._h.value for m in maps])
        X = (ctypes.c_void_p * n)(*x_ptrs)
        Y = (ctypes.c_void_p * n)(*y_ptrs)
        h = ctypes.c_void_p()
        rc = lib().ifgf_group_create(P, n, X, Y, ctypes.byref(h))
        if rc != IFGF_OK:
            raise IFGFError(f"ifgf_group_create failed ({rc}): {lib().ifgf_last_error().decode()}")
        self._h = h

    def launch(self, stream=None):
        rc = lib().ifgf_group_launch(self._h, ctypes.c_void_p(stream) if stream else None)
        if rc != IFGF_OK:
            raise IFGFError(lib().ifgf_last_error().decode())

    def stats(self):
        """Plans, classes, kernel launches and workgroups per group application."""
        s = np.empty(4, dtype=np.int64)
        rc = lib().ifgf_group_stats(self._h, s.ctypes.data_as(_lp))
        if rc != IFGF_OK:
            raise IFGFError(lib().ifgf_last_error().decode())
        return dict(zip(("plans", "classes", "launches", "workgroups"), s.tolist()))

    def synchronize(self):
        rc = lib().ifgf_group_synchronize(self._h)
        if rc != IFGF_OK:
            raise IFGFError(lib().ifgf_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().ifgf_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedPreconditioner:
    """The learned 1D FNO preconditioner (ifgf_fno_*): trained on the device on
    (right-hand side, density) pairs the solver produced, applied as z = M(v) on a
    ForwardMap's curve."""

    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        rc = lib().ifgf_fno_create(seed, ctypes.byref(h))
        if rc != IFGF_OK:
            raise IFGFError(f"ifgf_fno_create failed ({rc}): {lib().ifgf_fno_error(None).decode()}")
        self._h = h
        self.nparams = int(lib().ifgf_fno_num_params())

    def _check(self, rc, what):
        if rc != IFGF_OK:
            raise IFGFError(f"{what} failed ({rc}): {lib().ifgf_fno_error(self._h).decode()}")

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        self._check(lib().ifgf_fno_get_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise IFGFError("parameter vector of the wrong length")
        self._check(lib().ifgf_fno_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "set_params")

    def add_sample(self, fm, v, density):
        """Adds (v, density) on fm's curve; returns the sample index."""
        v, d = _c(v), _c(density)
        self._check(lib().ifgf_fno_add_sample(self._h, fm._h, _d(v.view(np.float64)), _d(d.view(np.float64))),
                    "add_sample")
        return int(lib().ifgf_fno_num_samples(self._h)) - 1

    @property
    def num_samples(self):
        return int(lib().ifgf_fno_num_samples(self._h))

    def clear_samples(self):
        self._check(lib().ifgf_fno_clear_samples(self._h), "clear_samples")

    def sample_data(self, s):
        n = ctypes.c_longlong()
        self._check(lib().ifgf_fno_sample_data(self._h, s, ctypes.byref(n), None, None, None, None), "sample_data")
        N = n.value
        X = np.empty((N, 11), dtype=np.float32)
        m = np.empty(N, dtype=np.float32)
        t, wt = np.empty(N), np.empty(N)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(lib().ifgf_fno_sample_data(self._h, s, ctypes.byref(n), X.ctypes.data_as(fp),
                                               m.ctypes.data_as(fp), _d(t), _d(wt)), "sample_data")
        return {"X": X, "mphys": m, "t": t, "wt": wt}

    def loss_grad(self, ids, grad=True):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        loss = ctypes.c_double()
        g = np.empty(self.nparams, dtype=np.float32) if grad else None
        self._check(lib().ifgf_fno_loss_grad(self._h, ids.ctypes.data_as(_ip), ids.size, ctypes.byref(loss),
                                             g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if grad else None),
                    "loss_grad")
        return loss.value, g

    def train(self, ids, epochs=10, batch=4, lr=1e-3, seed=0):
        """Adam over the samples ids; returns the mean loss of every epoch."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        hist = np.zeros(epochs)
        self._check(lib().ifgf_fno_train(self._h, ids.ctypes.data_as(_ip), ids.size, epochs, batch, lr, seed,
                                         _d(hist)), "train")
        return hist

    def apply(self, fm, v):
        """z = M(v) on fm's curve (host vectors)."""
        v = _c(v)
        z = np.empty_like(v)
        self._check(lib().ifgf_fno_apply(self._h, fm._h, v.ctypes.data, z.ctypes.data, 0, None), "apply")
        return z

    def apply_device(self, fm, v_ptr, z_ptr, stream=None):
        self._check(lib().ifgf_fno_apply(self._h, fm._h, ctypes.c_void_p(v_ptr), ctypes.c_void_p(z_ptr),
                                         IFGF_PTR_DEVICE, ctypes.c_void_p(stream) if stream else None), "apply")

    def close(self):
        if getattr(self, "_h", None):
            lib().ifgf_fno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fno_gemm_check(A, B, trans_a=False, trans_b=False):
    """The network's BF16 MFMA GEMM on host arrays: A B (A given transposed when
    trans_a, B when trans_b)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M, K = (A.shape[1], A.shape[0]) if trans_a else A.shape
    N = B.shape[0] if trans_b else B.shape[1]
    C = np.empty((M, N), dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = lib().ifgf_fno_gemm_check(M, N, K, int(trans_a), int(trans_b), A.ctypes.data_as(fp), B.ctypes.data_as(fp),
                                   C.ctypes.data_as(fp))
    if rc != IFGF_OK:
        raise IFGFError(lib().ifgf_fno_error(None).decode())
    return C


class LinearOperator:
    """LinearOperator(rows, cols, mapping) (ForwardMap/LinearOperator.hpp:27-60).

    ``op * x`` allocates the output and calls mapping(x, out) (LinearOperator.cpp:14-23)."""

    def __init__(self, rows, cols, mapping):
        self._rows, self._cols, self.mapping = int(rows), int(cols), mapping

    def rows(self):
        return self._rows

    def cols(self):
        return self._cols

    def __mul__(self, x):
        out = np.empty(self._rows, dtype=np.complex128)
        self.mapping(np.asarray(x, dtype=np.complex128), out)
        return out

    __matmul__ = __mul__


def gmres(op, rhs, restart=50, tol=1e-8, maxit=300, x0=None):
    """Householder GMRES on any LinearOperator (Eigen::opGMRES, opGMRES.hpp:68-269).

    Returns (x, iterations, estimated relative residual, converged)."""
    n = op.rows()
    b = _c(rhs)
    x = np.zeros(n, dtype=np.complex128) if x0 is None else _c(x0).copy()

    def cb(ctx, xp, yp):
        xin = np.ctypeslib.as_array(xp, shape=(2 * n,)).view(np.complex128)
        yout = np.ctypeslib.as_array(yp, shape=(2 * n,)).view(np.complex128)
        yout[:] = op * xin.copy()

    fn = APPLY_FN(cb)
    it, err = ctypes.c_int(), ctypes.c_double()
    rc = lib().ifgf_gmres_host(fn, None, n, _d(b.view(np.float64)), _d(x.view(np.float64)), restart,
                               tol, maxit, ctypes.byref(it), ctypes.byref(err))
    if rc not in (IFGF_OK, IFGF_ERR_NOT_CONVERGED):
        raise IFGFError(lib().ifgf_last_error().decode())
    return x, it.value, err.value, rc == IFGF_OK


def autotune_cone_order(curve, patch_lims, k, nlevels, tol, Np=10, formulation="combined", eta=None, delta=None,
                        p=4.0, precision="fp64", device_setup=True):
    """The cone-order autotuner (ifgf_autotune_cone_order): the cheapest (Ps, Pang)
    whose forward map is within tol of the unaccelerated operator. Returns
    (Ps, Pang, error, reached)."""
    lims = np.ascontiguousarray(patch_lims, dtype=np.float64)
    cp = np.ascontiguousarray(curve[1], dtype=np.float64)
    delta = default_delta(curve, lims) if delta is None else float(delta)
    surf = _Surface(curve[0], _d(cp), cp.size, _d(lims), lims.size - 1, Np, delta, float(p))
    flags = (IFGF_PLAN_DEVICE_SORT | IFGF_PLAN_DEVICE_PRECOMP | IFGF_PLAN_DEVICE_GEOMETRY
             | IFGF_PLAN_DEVICE_TREE) if device_setup else 0
    prm = _Params(_FORMS[formulation], _PREC[precision], float(k), float(k if eta is None else eta), 5, 5, nlevels, 1,
                  flags)
    a, b, e = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    rc = lib().ifgf_autotune_cone_order(ctypes.byref(surf), ctypes.byref(prm), tol, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(e))
    if rc not in (IFGF_OK, IFGF_ERR_NOT_CONVERGED):
        raise IFGFError(f"ifgf_autotune_cone_order failed ({rc}): {lib().ifgf_last_error().decode()}")
    return a.value, b.value, e.value, rc == IFGF_OK


def morton_sort(x, y, nlevels):
    """BoxTree::SortBox on the GPU: (sorting, leaf codes) of the points for an nlevels tree."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    s = np.empty(x.size, dtype=np.int64)
    c = np.empty(x.size, dtype=np.int64)
    rc = lib().ifgf_morton_sort(_d(x), _d(y), x.size, nlevels, s.ctypes.data_as(_lp), c.ctypes.data_as(_lp))
    if rc != IFGF_OK:
        raise IFGFError(f"ifgf_morton_sort failed ({rc}): {lib().ifgf_last_error().decode()}")
    return s, c


def hankel01(x):
    """H0^(1)(x), H1^(1)(x) evaluated by the GPU kernels' FP64 routine."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    h0 = np.empty(x.size, dtype=np.complex128)
    h1 = np.empty(x.size, dtype=np.complex128)
    rc = lib().ifgf_hankel01(_d(x), x.size, _d(h0.view(np.float64)), _d(h1.view(np.float64)))
    if rc != IFGF_OK:
        raise IFGFError(lib().ifgf_last_error().decode())
    return h0, h1
