"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is chosen so that a plausible slip (dropped term, wrong sign,
transposed index) fails it: mpmath golden vectors for the Hankel functions,
exactness identities for the quadrature, the reference's known answers
(circle single-layer eigenvalue, tests/ForwardMap/ComputeAx.cpp:69-70; IFGF
error against direct summation, IFGF_Source/main.cpp:104-148), the
double-layer / combined-field eigenvalues and the Mie series derived from the
addition theorem, and a direct solve for GMRES.
"""
import numpy as np
import pytest
import scipy.special as sp

import oracle as O


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ------------------------------------------------------------------ Hankel ---
def test_amos_hankel_matches_mpmath(hankel_golden):
    x = hankel_golden[:, 0]
    h0 = O.hankel(0, x)
    h1 = O.hankel(1, x)
    e0 = np.abs(h0 - (hankel_golden[:, 1] + 1j * hankel_golden[:, 2])) / np.abs(h0)
    e1 = np.abs(h1 - (hankel_golden[:, 3] + 1j * hankel_golden[:, 4])) / np.abs(h1)
    assert e0.max() < 5e-15 and e1.max() < 5e-15


def test_hankel_wronskian():
    x = np.logspace(-3, 3, 500)
    h0, h1 = O.hankel(0, x), O.hankel(1, x)
    # J1 Y0 - J0 Y1 = 2 / (pi x)
    w = h1.real * h0.imag - h0.real * h1.imag
    assert np.max(np.abs(w * np.pi * x / 2 - 1)) < 1e-12


# -------------------------------------------------------------- quadrature ---
@pytest.mark.parametrize("n", [5, 10, 20, 40])
def test_fejer_exact_on_polynomials(n):
    x, w = O.fejer(n)
    assert np.allclose(x, np.cos(np.pi * (2 * np.arange(n) + 1) / (2 * n)), atol=0, rtol=1e-15)
    assert abs(w.sum() - 2.0) < 1e-14
    for d in range(n):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.dot(w, x ** d) - exact) < 1e-13


@pytest.mark.parametrize("which,alpha", [(0, 0.3), (0, -0.77), (-1, -1.0), (1, 1.0)])
def test_rp_change_of_variables(which, alpha):
    t, w = O.fejer(40)
    xi, dxi = O.rp_cov(which, t, alpha, 6.0)
    # maps [-1, 1] onto [-1, 1] (integral of the Jacobian is the length 2)
    assert np.all(xi >= -1) and np.all(xi <= 1)
    assert abs(np.dot(w, dxi) - 2.0) < 1e-6  # the Jacobian is only piecewise smooth at t = 0
    # integrates a log singularity at alpha: int_{-1}^{1} log|x - alpha| dx
    a = alpha
    exact = ((1 - a) * np.log(abs(1 - a)) if a != 1 else 0) + ((1 + a) * np.log(abs(1 + a)) if a != -1 else 0) - 2
    d = np.abs(xi - a)
    # nodes that round onto the singularity contribute 0, as GF does for r < 1e-14
    approx = np.dot(w * dxi, np.where(d > 0, np.log(np.where(d > 0, d, 1.0)), 0.0))
    assert abs(approx - exact) < 1e-5


# ------------------------------------------------------- forward map, circle ---
def circle_fm(npatch, k, F, nlev, Ps=5, Pang=5, p=6.0, delta=None):
    lims = np.linspace(0, 2 * np.pi, npatch + 1)
    if delta is None:
        delta = 0.75 * 2 * np.sin(np.pi / npatch)
    return O.ForwardMap(O.CIRCLE, [0.0, 0.0, 1.0], lims, Np=10, formulation=F, k=k, eta=k,
                        delta=delta, p=p, Ps=Ps, Pang=Pang, nlevels=nlev)


def circle_eig(F, m, k):
    lam = 1j * np.pi / 2 * sp.jv(m, k) * sp.hankel1(m, k)
    mu = 1j * np.pi * k / 4 * (sp.jvp(m, k) * sp.hankel1(m, k) + sp.jv(m, k) * sp.h1vp(m, k))
    return {O.SINGLE_LAYER: lam, O.DOUBLE_LAYER: 0.5 + mu, O.COMBINED: 0.5 + mu - 1j * k * lam}[F]


def test_compute_ax_known_answer():
    """tests/ForwardMap/ComputeAx.cpp: unit circle, k = 100 pi, one wavelength
    per patch (Curve::compute_patch_lims), Np = 10, delta = 0.01, nlevels 5,
    SingleLayer applied to exp(3 i theta) equals lambda_3 exp(3 i theta)."""
    k = 100 * np.pi
    lims = O.patch_lims(O.CIRCLE, [0.0, 0.0, 1.0], 1.0, k)
    assert lims.size - 1 == 315
    fm = O.ForwardMap(O.CIRCLE, [0.0, 0.0, 1.0], lims, Np=10, formulation=O.SINGLE_LAYER, k=k,
                      eta=0.0, delta=0.01, p=4.0, nlevels=5)
    x, y, *_ = fm.points()
    u = np.exp(3j * np.arctan2(y, x))
    lam = circle_eig(O.SINGLE_LAYER, 3, k)
    ya = fm.apply(u)
    assert rel(ya, lam * u) < 2e-3      # IFGF 5x5 interpolation error level
    yd = fm.apply(u, accelerated=False)
    assert rel(yd, lam * u) < 1e-4      # quadrature + RP correction


@pytest.mark.parametrize("F", [O.SINGLE_LAYER, O.DOUBLE_LAYER, O.COMBINED])
def test_circle_eigenvalues_direct(F):
    fm = circle_fm(40, 10.0, F, 3)
    x, y, *_ = fm.points()
    for m in (0, 3, 8):
        u = np.exp(1j * m * np.arctan2(y, x))
        assert rel(fm.apply(u, accelerated=False), circle_eig(F, m, 10.0) * u) < 1e-6


def test_circle_eigenvalue_sign_check():
    """A dropped 1/2 identity or a flipped double-layer sign must fail."""
    fm = circle_fm(40, 10.0, O.COMBINED, 3)
    x, y, *_ = fm.points()
    u = np.exp(2j * np.arctan2(y, x))
    y_ = fm.apply(u, accelerated=False)
    lam = circle_eig(O.SINGLE_LAYER, 2, 10.0)
    mu = circle_eig(O.DOUBLE_LAYER, 2, 10.0) - 0.5
    assert rel(y_, (0.5 + mu - 10j * lam) * u) < 1e-6
    assert rel(y_, (-0.5 + mu - 10j * lam) * u) > 0.1
    assert rel(y_, (0.5 - mu - 10j * lam) * u) > 0.1


def test_ifgf_error_levels_and_convergence_in_p():
    """IFGF vs direct summation (IFGF_Source/main.cpp:104-148): the 5x5 error
    sits at the reference's level (~1e-4..1e-3) and drops with more nodes."""
    errs = []
    for P in (5, 7, 9):
        fm = circle_fm(40, 10.0, O.SINGLE_LAYER, 3, Ps=P, Pang=P)
        x, y, *_ = fm.points()
        rng = np.random.default_rng(0)
        u = rng.standard_normal(fm.N) + 1j * rng.standard_normal(fm.N)
        errs.append(rel(fm.apply(u), fm.apply(u, accelerated=False)))
    assert 1e-4 < errs[0] < 3e-3
    assert errs[1] < errs[0] / 5 and errs[2] < errs[1] / 5


def test_boxtree_standalone_vs_direct():
    """BoxTree::Solve on a bare point set against the direct sum that skips i == j."""
    n = 2000
    th = 2 * np.pi * np.arange(n) / n
    x, y = np.cos(th), np.sin(th)
    rng = np.random.default_rng(1)
    d = rng.uniform(0, 1, n) * (1 + 1j)
    out, st = O.boxtree_solve(x, y, x, y, d, nlevels=6, k=60.0)
    tg = np.arange(0, n, 7)
    ref = O.direct(x, y, x, y, d, 60.0, tg)
    assert 1e-5 < rel(out[tg], ref) < 2e-3
    assert st["leaf"] > 0 and st["interp"] > 0 and st["prop"] > 0


def test_field_solve_reduces_to_boxtree_at_the_points():
    """The separate-target BoxTree (no reference counterpart) with the targets
    equal to the points is the reference's BoxTree::Solve: same tree, same
    cones, same sums (the self pair it adds has a zero kernel)."""
    n = 1200
    th = 2 * np.pi * (np.arange(n) + 0.25) / n
    x, y = np.cos(th) + 0.3 * np.cos(2 * th), 1.2 * np.sin(th)
    rng = np.random.default_rng(2)
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for F, eta in ((O.SINGLE_LAYER, 0.0), (O.COMBINED, 25.0)):
        a, sa = O.boxtree_solve(x, y, np.cos(th), np.sin(th), d, nlevels=5, k=25.0, formulation=F, eta=eta)
        b, sb = O.field_solve(x, y, np.cos(th), np.sin(th), d, x, y, nlevels=5, k=25.0, formulation=F,
                              eta=eta)
        assert np.array_equal(a, b)
        assert sb["near"] == sa["near"] + n
        for key in ("leaf", "prop", "interp", "cones"):
            assert sa[key] == sb[key], key


@pytest.mark.parametrize("F", [O.SINGLE_LAYER, O.DOUBLE_LAYER, O.COMBINED])
def test_field_solve_converges_to_propagate_unacc(F):
    """IFGF at separate targets (a ring outside the circle and a grid through
    it) against the direct field sum of propagate_unacc (ForwardMap.tpp:702-727):
    the error falls geometrically with the interpolation order."""
    om = circle_fm(80, 10.0, F, 4)
    x, y, nx, ny, w = om.points()
    rng = np.random.default_rng(3)
    u = rng.standard_normal(om.N) + 1j * rng.standard_normal(om.N)
    tt = rng.uniform(0, 2 * np.pi, 200)
    gx, gy = np.meshgrid(np.linspace(-2.0, 2.0, 17), np.linspace(-2.0, 2.0, 17))
    tx = np.concatenate([1.5 * np.cos(tt), gx.ravel()])
    ty = np.concatenate([1.5 * np.sin(tt), gy.ravel()])
    ref = om.propagate(tx, ty, u)
    errs = []
    for P in (5, 7, 9):
        out, st = O.field_solve(x, y, nx, ny, u * w, tx, ty, nlevels=5, k=10.0, formulation=F, eta=10.0,
                                Ps=P, Pang=P)
        assert st["near"] > 0 and st["interp"] > 0 and st["prop"] > 0
        errs.append(rel(out, ref))
    assert 1e-5 < errs[0] < 1e-2
    assert errs[1] < errs[0] / 5 and errs[2] < errs[1] / 5


def test_gmres_dense_matches_direct_solve():
    rng = np.random.default_rng(2)
    n = 60
    A = np.eye(n) * 3 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, it, err, rc = O.gmres_dense(A, b, restart=20, tol=1e-12, maxit=500)
    assert rc == 0 and err < 1e-12
    assert rel(x, np.linalg.solve(A, b)) < 1e-10


def mie(r, th, k, d_angle=0.0, M=None):
    M = M or int(k + 40)
    m = np.arange(-M, M + 1)
    c = -(1j ** m) * sp.jv(m, k) / sp.hankel1(m, k)
    return np.sum(c[None, :] * sp.hankel1(m[None, :], k * r) * np.exp(1j * m[None, :] * (th[:, None] - d_angle)), axis=1)


def test_mie_series_solve():
    """Sound-soft circle, k = 10, plane wave d = (1, 0): CFIE solve by the
    reference's GMRES, then the scattered field at 64 points on r = 3."""
    fm = circle_fm(40, 10.0, O.COMBINED, 3, Ps=9, Pang=9)
    x, y, *_ = fm.points()
    rhs = -np.exp(1j * 10.0 * x)
    dens, it, err, hist = fm.gmres(rhs, restart=50, tol=1e-8)
    assert err < 1e-8 and np.all(np.diff(hist) <= 1e-15)
    th = 2 * np.pi * np.arange(64) / 64
    us = fm.propagate(3 * np.cos(th), 3 * np.sin(th), dens)
    assert rel(us, mie(3.0, th, 10.0)) < 1e-5


def test_gmres_is_residual_minimising():
    """The Householder GMRES residual after k steps equals the minimum over the
    Krylov space (computed by Arnoldi + least squares), for complex data."""
    rng = np.random.default_rng(5)
    n = 120
    A = np.eye(n) + 0.9 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    V = np.zeros((n, 41), complex)
    H = np.zeros((41, 40), complex)
    beta = np.linalg.norm(b)
    V[:, 0] = b / beta
    best = []
    for j in range(40):
        w = A @ V[:, j]
        for _ in range(2):
            for i in range(j + 1):
                c = np.vdot(V[:, i], w)
                H[i, j] += c
                w = w - c * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
        e = np.zeros(j + 2, complex)
        e[0] = beta
        y = np.linalg.lstsq(H[: j + 2, : j + 1], e, rcond=None)[0]
        best.append(np.linalg.norm(e - H[: j + 2, : j + 1] @ y) / beta)
    for k in (1, 2, 5, 10, 40):
        x, it, err, rc = O.gmres_dense(A, b, restart=40, tol=1e-14, maxit=k)
        true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert abs(true - best[k - 1]) < 1e-10 * max(1.0, best[k - 1])
        assert abs(err - true) < 1e-10


# ------------------------------------------------- benchmark input generator ---
def test_seeded_density_generator_matches_fixture():
    """The benchmark's counter-based generator (splitmix64 + Box-Muller), in the
    oracle's C++ and in ifgf2d_amd.configs, against the fixture written from an
    independent Python-integer / mpmath restatement, and splitmix64's published
    first output for seed 0 (0xE220A8397B1DCDAF)."""
    import json
    import os

    from ifgf2d_amd import configs

    with open(os.path.join(os.path.dirname(__file__), "golden", "seeded_density.json")) as f:
        g = json.load(f)
    assert int(g["splitmix64"]["0"][0]) == 0xE220A8397B1DCDAF
    for s, vals in g["splitmix64"].items():
        want = np.array([int(v) for v in vals], dtype=np.uint64)
        assert np.array_equal(O.splitmix64(int(s), 16), want)
        assert np.array_equal(configs.splitmix64(int(s), 16), want)
        assert np.array_equal(configs.splitmix64(int(s), 10, start=6), want[6:])
    # FP64 evaluation of the 30-digit values: the angle 2 pi u2 carries ~eps * 2 pi,
    # times the radius
    want = np.array([complex(a, b) for a, b in g["density_seed0"]])
    for got in (O.seeded_density(8), configs.seeded_density(8)):
        assert np.all(np.abs(got - want) <= 2e-15 * np.maximum(1.0, np.abs(want)))
    # C++ and numpy agree to the math libraries' last bits; column c of an n-entry
    # block continues the stream at entry c n
    a, b = O.seeded_density(5000, seed=3, column=2), configs.seeded_density(5000, seed=3, column=2)
    assert np.all(np.abs(a - b) <= 2e-15 * np.maximum(1.0, np.abs(a)))
    assert np.array_equal(configs.seeded_density(10, column=1), configs.seeded_density(20)[10:])
    z = configs.seeded_density(200000, seed=1)
    assert abs(z.real.mean()) < 0.01 and abs(z.imag.mean()) < 0.01
    assert abs(z.real.var() - 1) < 0.02 and abs(z.imag.var() - 1) < 0.02
    assert abs(np.mean(z.real * z.imag)) < 0.01


def test_oracle_threads_do_not_change_results():
    """The oracle's OpenMP loops sum every output in the serial order: one
    thread and all threads give the same bits (kite, N = 2,000, 5 levels)."""
    import ifgf2d_amd as I
    from ifgf2d_amd import configs
    lims = I.equal_parameter_lims(200)
    curve = I.kite()
    n0 = O.set_threads(0)
    try:
        O.set_threads(1)
        om = O.ForwardMap(curve[0], curve[1], lims, k=20.0, delta=I.default_delta(curve, lims), p=6.0,
                          nlevels=5)
        x = configs.seeded_density(om.N)
        y1 = om.apply(x)
        pts1 = om.precomputation(7, 10 + om.near_singular()[0][7])
        O.set_threads(max(2, n0))
        om2 = O.ForwardMap(curve[0], curve[1], lims, k=20.0, delta=I.default_delta(curve, lims), p=6.0,
                           nlevels=5)
        assert np.array_equal(om2.apply(x), y1)
        assert np.array_equal(om2.precomputation(7, 10 + om2.near_singular()[0][7]), pts1)
    finally:
        O.set_threads(n0)
