"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SINGLE_LAYER, DOUBLE_LAYER, COMBINED = 0, 1, 2
CIRCLE, KITE, STAR = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_longlong)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _LIB = ctypes.CDLL(path)
        L = _LIB
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_hankel.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_longlong]
        L.oracle_splitmix.argtypes = [ctypes.c_ulonglong, ctypes.c_longlong, ctypes.c_longlong,
                                      ctypes.POINTER(ctypes.c_ulonglong)]
        L.oracle_seeded_density.argtypes = [ctypes.c_ulonglong, ctypes.c_longlong, ctypes.c_int, _dp]
        L.oracle_fejer.argtypes = [ctypes.c_int, _dp, _dp]
        L.oracle_rp_cov.argtypes = [ctypes.c_int, _dp, ctypes.c_longlong, ctypes.c_double,
                                    ctypes.c_double, _dp, _dp]
        L.oracle_patch_lims.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _dp, ctypes.c_int]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = ctypes.c_int
        L.oracle_curve_eval.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_longlong, _dp, _dp,
                                        _dp, _dp]
        L.oracle_curve_diff.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, ctypes.c_longlong,
                                        _dp, _dp]
        L.oracle_fm_create.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_fm_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_fm_size.argtypes = [ctypes.c_void_p]
        L.oracle_fm_size.restype = ctypes.c_longlong
        L.oracle_fm_points.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.oracle_fm_ns.argtypes = [ctypes.c_void_p, _ip, _lp, _dp]
        L.oracle_fm_ns.restype = ctypes.c_longlong
        L.oracle_fm_precomp.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.oracle_fm_apply.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.oracle_fm_stats.argtypes = [ctypes.c_void_p, _lp]
        L.oracle_fm_sorting.argtypes = [ctypes.c_void_p, _lp]
        L.oracle_fm_level_cones.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _lp]
        L.oracle_fm_level_cones.restype = ctypes.c_longlong
        L.oracle_fm_propagate.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_longlong, _dp, _dp]
        L.oracle_fm_gmres.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, _ip, _dp, _dp, ctypes.c_int]
        L.oracle_gmres_dense.argtypes = [_dp, ctypes.c_longlong, _dp, _dp, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, _ip, _dp]
        L.oracle_boxtree_solve.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_longlong, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_int, _dp, _dp, _lp]
        L.oracle_field_solve.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_longlong, _dp, _dp,
                                         ctypes.c_longlong, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         _dp, _dp, _lp]
        L.oracle_direct.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_longlong, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_double, _dp, _lp, ctypes.c_longlong,
                                    _dp]
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(a):
    return a.ctypes.data_as(_lp)


def _err():
    return lib().oracle_last_error().decode()


def hankel(n, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(2 * x.size)
    lib().oracle_hankel(n, _d(x), _d(out), x.size)
    return out.view(np.complex128)


def fejer(n):
    a, b = np.empty(n), np.empty(n)
    lib().oracle_fejer(n, _d(a), _d(b))
    return a, b


def rp_cov(which, t, alpha, p):
    t = np.ascontiguousarray(t, dtype=np.float64)
    xi, dxi = np.empty_like(t), np.empty_like(t)
    lib().oracle_rp_cov(which, _d(t), t.size, alpha, p, _d(xi), _d(dxi))
    return xi, dxi


def set_threads(n=0):
    """Threads of the oracle's OpenMP loops (0: the runtime default); returns
    the count in effect. Results do not depend on it."""
    return lib().oracle_set_threads(int(n))


def curve_eval(curve_type, params, t):
    """(x, y, x', y') of the oracle's Curve at the parameters t."""
    prm = np.ascontiguousarray(params, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = [np.empty_like(t) for _ in range(4)]
    if lib().oracle_curve_eval(curve_type, _d(prm), prm.size, _d(t), t.size, *[_d(o) for o in out]):
        raise RuntimeError(_err())
    return tuple(out)


def curve_diff(curve_type, params, a, d):
    """(x(a) - x(a - d), y(a) - y(a - d)) by the oracle's Curve::diff."""
    prm = np.ascontiguousarray(params, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    dx, dy = np.empty_like(a), np.empty_like(a)
    if lib().oracle_curve_diff(curve_type, _d(prm), prm.size, _d(a), _d(d), a.size, _d(dx), _d(dy)):
        raise RuntimeError(_err())
    return dx, dy


def patch_lims(curve_type, params, wavelengths_per_patch, k):
    prm = np.ascontiguousarray(params, dtype=np.float64)
    out = np.empty(1 << 20)
    n = lib().oracle_patch_lims(curve_type, _d(prm), prm.size, wavelengths_per_patch, k, _d(out),
                                out.size)
    if n <= 0:
        raise RuntimeError(_err())
    return out[:n].copy()


class ForwardMap:
    """Oracle forward map on one closed curve (ForwardMap<Np, F, Ps, Pang>)."""

    def __init__(self, curve_type, params, lims, Np=10, formulation=COMBINED, k=10.0, eta=None,
                 delta=None, p=4.0, Ps=5, Pang=5, nlevels=3):
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.lims = np.ascontiguousarray(lims, dtype=np.float64)
        eta = k if eta is None else eta
        if delta is None:
            raise ValueError("delta is required")
        h = ctypes.c_void_p()
        rc = lib().oracle_fm_create(curve_type, _d(self.params), self.params.size, _d(self.lims),
                                    self.lims.size - 1, Np, formulation, k, eta, delta, p, Ps,
                                    Pang, nlevels, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("oracle ForwardMap: " + _err())
        self.h = h
        self.N = lib().oracle_fm_size(h)
        self.Np = Np
        self.npatches = self.lims.size - 1

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_fm_destroy(self.h)
            self.h = None

    def points(self):
        x, y, nx, ny, w = (np.empty(self.N) for _ in range(5))
        lib().oracle_fm_points(self.h, _d(x), _d(y), _d(nx), _d(ny), _d(w))
        return x, y, nx, ny, w

    def near_singular(self):
        counts = np.empty(self.npatches, dtype=np.int32)
        tot = lib().oracle_fm_ns(self.h, counts.ctypes.data_as(_ip), None, None)
        idx = np.empty(max(tot, 1), dtype=np.int64)
        ts = np.empty(max(tot, 1))
        lib().oracle_fm_ns(self.h, counts.ctypes.data_as(_ip), _l(idx), _d(ts))
        return counts, idx[:tot], ts[:tot]

    def precomputation(self, patch, ncols):
        out = np.empty(2 * self.Np * ncols)
        lib().oracle_fm_precomp(self.h, patch, _d(out))
        return out.view(np.complex128).reshape(ncols, self.Np)  # [column][cheb degree]

    def apply(self, x, accelerated=True):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        rc = lib().oracle_fm_apply(self.h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                   1 if accelerated else 0)
        if rc != 0:
            raise RuntimeError(_err())
        return y

    def stats(self):
        s = np.empty(5, dtype=np.int64)
        lib().oracle_fm_stats(self.h, _l(s))
        return dict(zip(("near", "leaf", "prop", "interp", "cones"), s.tolist()))

    def sorting(self):
        s = np.empty(self.N, dtype=np.int64)
        lib().oracle_fm_sorting(self.h, _l(s))
        return s

    def level_cones(self, level):
        n = lib().oracle_fm_level_cones(self.h, level, None, None)
        b, q = np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int64)
        lib().oracle_fm_level_cones(self.h, level, _l(b), _l(q))
        return b, q

    def propagate(self, tx, ty, dens):
        tx = np.ascontiguousarray(tx, dtype=np.float64)
        ty = np.ascontiguousarray(ty, dtype=np.float64)
        d = np.ascontiguousarray(dens, dtype=np.complex128)
        out = np.empty(tx.size, dtype=np.complex128)
        lib().oracle_fm_propagate(self.h, _d(tx), _d(ty), tx.size, d.ctypes.data_as(_dp),
                                  out.ctypes.data_as(_dp))
        return out

    def gmres(self, rhs, restart=50, tol=1e-8, maxit=1000, x0=None):
        b = np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.complex128)
        it, err = ctypes.c_int(), ctypes.c_double()
        hist = np.zeros(maxit + 1)
        rc = lib().oracle_fm_gmres(self.h, b.ctypes.data_as(_dp), x.ctypes.data_as(_dp), restart,
                                   tol, maxit, ctypes.byref(it), ctypes.byref(err), _d(hist),
                                   hist.size)
        if rc == 1:
            raise RuntimeError(_err())
        return x, it.value, err.value, hist[: it.value]


def gmres_dense(A, b, restart=50, tol=1e-12, maxit=1000):
    A = np.asfortranarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros_like(b)
    it, err = ctypes.c_int(), ctypes.c_double()
    rc = lib().oracle_gmres_dense(A.ctypes.data_as(_dp), b.size, b.ctypes.data_as(_dp),
                                  x.ctypes.data_as(_dp), restart, tol, maxit, ctypes.byref(it),
                                  ctypes.byref(err))
    return x, it.value, err.value, rc


def boxtree_solve(x, y, nx, ny, dens, nlevels, k, formulation=SINGLE_LAYER, eta=0.0, Ps=5, Pang=5):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, nx, ny)]
    d = np.ascontiguousarray(dens, dtype=np.complex128)
    out = np.empty_like(d)
    st = np.empty(5, dtype=np.int64)
    rc = lib().oracle_boxtree_solve(*[_d(a) for a in arrs], arrs[0].size, nlevels, k, formulation,
                                    eta, Ps, Pang, d.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                    _l(st))
    if rc != 0:
        raise RuntimeError(_err())
    return out, dict(zip(("near", "leaf", "prop", "interp", "cones"), st.tolist()))


def field_solve(x, y, nx, ny, dens, tx, ty, nlevels, k, formulation=SINGLE_LAYER, eta=0.0, Ps=5,
                Pang=5):
    """BoxTree solve with a separate target set (tx, ty): the field of the
    intensities dens at the targets, near field included (coincident points
    contribute zero). Returns (values in target order, counts)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, nx, ny)]
    t = [np.ascontiguousarray(a, dtype=np.float64) for a in (tx, ty)]
    d = np.ascontiguousarray(dens, dtype=np.complex128)
    out = np.empty(t[0].size, dtype=np.complex128)
    st = np.empty(5, dtype=np.int64)
    rc = lib().oracle_field_solve(*[_d(a) for a in arrs], arrs[0].size, _d(t[0]), _d(t[1]), t[0].size,
                                  nlevels, k, formulation, eta, Ps, Pang, d.ctypes.data_as(_dp),
                                  out.ctypes.data_as(_dp), _l(st))
    if rc != 0:
        raise RuntimeError(_err())
    return out, dict(zip(("near", "leaf", "prop", "interp", "cones"), st.tolist()))


def direct(x, y, nx, ny, dens, k, targets, formulation=SINGLE_LAYER, eta=0.0):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, nx, ny)]
    d = np.ascontiguousarray(dens, dtype=np.complex128)
    t = np.ascontiguousarray(targets, dtype=np.int64)
    out = np.empty(t.size, dtype=np.complex128)
    lib().oracle_direct(*[_d(a) for a in arrs], arrs[0].size, k, formulation, eta,
                        d.ctypes.data_as(_dp), _l(t), t.size, out.ctypes.data_as(_dp))
    return out


def splitmix64(seed, n, start=0):
    out = np.empty(n, dtype=np.uint64)
    lib().oracle_splitmix(seed, n, start, out.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)))
    return out


def seeded_density(n, seed=0, column=0):
    out = np.empty(n, dtype=np.complex128)
    lib().oracle_seeded_density(seed, n, column, out.ctypes.data_as(_dp))
    return out
