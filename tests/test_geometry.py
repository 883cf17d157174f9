"""Pins the curve geometry of the oracle and of the product on the non-circular
curves (kite, C3 star, random stars), where the circle's closed forms do not
reach.

The curves are defined by their positions only (Geometry/Curve.hpp:106-146:
kite x = A cos t + B cos 2t + C, y = D sin t; star r(t) = R (1 + sum a_m cos mt
+ b_m sin mt)). Everything derived from them is checked against mpmath:
  * x'(t), y'(t) against mpmath's numerical derivative of the positions (a
    dropped factor 2 in -2B sin 2t, or a wrong sign in the star's r'(t), fails);
  * Curve::diff(a, d) = (x(a) - x(a - d), y(a) - y(a - d)) against the
    subtraction carried out in 40-digit arithmetic, down to d = 1e-9;
  * the quadrature weights (Fejer weight x Jacobian x |x'|, compute_intensities,
    ForwardMap.tpp:425-435) against mpmath.quad of the arc length, the first
    moments and, with the normals, the enclosed area (divergence theorem);
  * the normals: unit length, orthogonal to mpmath's tangent, outward;
  * the whole pipeline on the kite and a star by a closed form: for a point
    source x0 inside the obstacle, u = H0(k |x - x0|) is the radiating
    exterior solution with Dirichlet data g = u on the curve, so the CFIE
    solution's field outside must reproduce it.
"""
import mpmath as mp
import numpy as np
import pytest
import scipy.special as sp

import ifgf2d_amd as I
import oracle as O
from ifgf2d_amd import configs

mp.mp.dps = 40


def _curves():
    rs = configs.random_stars(1, k=100.0, seed=0)[0]["curve"]
    return {
        "kite": I.kite(),
        "star_c3": I.star(a=[0, 0, 0, 0, 0, 0.3]),
        "random_star": rs,
        "offset_star": I.star(cx=0.3, cy=-0.2, R=1.7, a=[0.05, 0.0, 0.1, -0.04], b=[0.0, 0.08, -0.03, 0.02]),
        "circle": I.circle(0.1, -0.4, 1.3),
    }


CURVES = _curves()


def mp_position(curve):
    """The curve's definition x(t), y(t) in mpmath."""
    typ, p = curve
    if typ == 0:
        cx, cy, R = (mp.mpf(v) for v in p[:3])
        return lambda t: cx + R * mp.cos(t), lambda t: cy + R * mp.sin(t)
    if typ == 1:
        A, B, C, D = (mp.mpf(v) for v in p[:4])
        return lambda t: A * mp.cos(t) + B * mp.cos(2 * t) + C, lambda t: D * mp.sin(t)
    cx, cy, R = (mp.mpf(v) for v in p[:3])
    n = int(p[3])
    a = [mp.mpf(v) for v in p[4:4 + n]]
    b = [mp.mpf(v) for v in p[4 + n:4 + 2 * n]]

    def r(t):
        return R * (1 + mp.fsum(a[m] * mp.cos(m * t) + b[m] * mp.sin(m * t) for m in range(n)))
    return lambda t: cx + r(t) * mp.cos(t), lambda t: cy + r(t) * mp.sin(t)


def _ts(n=100, seed=3):
    return np.sort(np.random.default_rng(seed).uniform(0.0, 2 * np.pi, n))


@pytest.mark.parametrize("name", list(CURVES))
def test_positions_and_derivatives_match_mpmath(name):
    curve = CURVES[name]
    X, Y = mp_position(curve)
    t = _ts()
    ex = np.array([float(X(v)) for v in t])
    ey = np.array([float(Y(v)) for v in t])
    exp_ = np.array([float(mp.diff(X, mp.mpf(v))) for v in t])
    eyp = np.array([float(mp.diff(Y, mp.mpf(v))) for v in t])
    scale = max(np.abs(exp_).max(), np.abs(eyp).max())
    for src, (x, y, xp, yp) in (("product", I.curve_eval(curve, t)), ("oracle", O.curve_eval(curve[0], curve[1], t))):
        assert np.abs(x - ex).max() < 1e-14 * max(1.0, np.abs(ex).max()), src
        assert np.abs(y - ey).max() < 1e-14 * max(1.0, np.abs(ey).max()), src
        assert np.abs(xp - exp_).max() < 1e-13 * scale, (src, np.abs(xp - exp_).max())
        assert np.abs(yp - eyp).max() < 1e-13 * scale, (src, np.abs(yp - eyp).max())


@pytest.mark.parametrize("name", list(CURVES))
@pytest.mark.parametrize("d", [1e-9, 1e-6, 1e-3, 0.1, 1.0, 3.0])
def test_curve_diff_matches_extended_precision_subtraction(name, d):
    curve = CURVES[name]
    X, Y = mp_position(curve)
    a = _ts(60, seed=11)
    dd = np.full_like(a, d)
    ex = np.array([float(X(mp.mpf(v)) - X(mp.mpf(v) - mp.mpf(d))) for v in a])
    ey = np.array([float(Y(mp.mpf(v)) - Y(mp.mpf(v) - mp.mpf(d))) for v in a])
    ref = np.hypot(ex, ey)
    for src, (dx, dy) in (("product", I.curve_diff(curve, a, dd)),
                          ("oracle", O.curve_diff(curve[0], curve[1], a, dd))):
        err = np.hypot(dx - ex, dy - ey) / ref
        assert err.max() < 1e-13, (src, err.max())
        # the plain subtraction loses relative accuracy as d shrinks; diff does not
    if d <= 1e-6:
        x, y, _, _ = I.curve_eval(curve, a)
        x2, y2, _, _ = I.curve_eval(curve, a - d)
        naive = np.hypot(x - x2 - ex, y - y2 - ey) / ref
        assert naive.max() > 1e-11


def _mp_curve_integrals(curve):
    """Perimeter, int x ds, int y^2 ds and the enclosed area by mpmath.quad."""
    X, Y = mp_position(curve)

    def speed(t):
        return mp.sqrt(mp.diff(X, t) ** 2 + mp.diff(Y, t) ** 2)
    # split [0, 2 pi] so the quadrature resolves the star's oscillations
    pts = mp.linspace(0, 2 * mp.pi, 33)
    old = mp.mp.dps
    mp.mp.dps = 20
    try:
        L = mp.quad(speed, pts)
        mx = mp.quad(lambda t: X(t) * speed(t), pts)
        my2 = mp.quad(lambda t: Y(t) ** 2 * speed(t), pts)
        area = mp.quad(lambda t: (X(t) * mp.diff(Y, t) - Y(t) * mp.diff(X, t)) / 2, pts)
    finally:
        mp.mp.dps = old
    return float(L), float(mx), float(my2), float(area)


@pytest.mark.parametrize("name", ["kite", "star_c3", "random_star", "offset_star"])
def test_weights_and_normals_integrate_the_curve(name):
    """Sum w f(x_i) against mpmath's line integrals of f, for the product's
    plan (host-only) and the oracle: pins the weights, Jacobians and normals."""
    curve = CURVES[name]
    L, mx, my2, area = _mp_curve_integrals(curve)
    npatch = 200
    lims = I.equal_parameter_lims(npatch)
    fm = I.ForwardMap(curve, lims, k=5.0, nlevels=4, host_only=True, p=6.0)
    om = O.ForwardMap(curve[0], curve[1], lims, k=5.0, delta=fm.delta, p=6.0, nlevels=4)
    for src, (x, y, nx, ny, w) in (("product", fm.points()), ("oracle", om.points())):
        assert abs(w.sum() - L) < 1e-12 * L, (src, w.sum(), L)
        assert abs(np.dot(w, x) - mx) < 1e-12 * L * np.abs(x).max(), src
        assert abs(np.dot(w, y * y) - my2) < 1e-12 * L * (y * y).max(), src
        # divergence theorem: int x n_x ds = int y n_y ds = area (outward normals)
        assert abs(np.dot(w, x * nx) - area) < 1e-11 * abs(area), (src, np.dot(w, x * nx), area)
        assert abs(np.dot(w, y * ny) - area) < 1e-11 * abs(area), src
        assert np.abs(np.hypot(nx, ny) - 1).max() < 1e-15
    # normals orthogonal to the tangent from mpmath at the nodes' parameters
    X, Y = mp_position(curve)
    nodes = np.cos(np.pi * (2 * np.arange(10) + 1) / 20)
    x, y, nx, ny, w = fm.points()
    for p_ in range(0, npatch, 37):
        a, b = lims[p_], lims[p_ + 1]
        for j in (0, 4, 9):
            t = mp.mpf((b - a) / 2.0 * (nodes[j] + 1.0) + a)
            tx, ty = float(mp.diff(X, t)), float(mp.diff(Y, t))
            i = p_ * 10 + j
            assert abs(nx[i] * tx + ny[i] * ty) < 1e-13 * np.hypot(tx, ty)
            assert abs(x[i] - float(X(t))) < 1e-14 * max(1.0, abs(x[i]))


def exterior_point_source(x, y, x0, k):
    return sp.hankel1(0, k * np.hypot(x - x0[0], y - x0[1]))


@pytest.mark.parametrize("name,x0", [("kite", (-0.2, 0.3)), ("star_c3", (0.15, -0.25)),
                                     ("offset_star", (0.5, 0.1))])
def test_oracle_point_source_closed_form(name, x0):
    """The oracle's whole CFIE pipeline on a non-circular curve: Dirichlet data
    H0(k |x - x0|) from an interior source, solve, field at 32 points on
    r = 3 equals the source's field (64 patches, 9 x 9 cones)."""
    curve = CURVES[name]
    k = 8.0
    lims = I.equal_parameter_lims(64)
    delta = I.default_delta(curve, lims)
    om = O.ForwardMap(curve[0], curve[1], lims, k=k, delta=delta, p=6.0, nlevels=4, Ps=9, Pang=9)
    x, y, *_ = om.points()
    dens, it, err, _ = om.gmres(exterior_point_source(x, y, x0, k), restart=60, tol=1e-12)
    th = 2 * np.pi * np.arange(32) / 32
    tx, ty = 3.0 * np.cos(th), 3.0 * np.sin(th)
    u = om.propagate(tx, ty, dens)
    ue = exterior_point_source(tx, ty, x0, k)
    e = np.linalg.norm(u - ue) / np.linalg.norm(ue)
    print(f"{name}: {it} iterations, field error {e:.2e}")
    # measured 1.1e-7 (kite), 1.1e-6 (C3 star: the 9 x 9 interpolation error), 9.8e-8
    assert e < 3e-6
    # the check is sharp: a density perturbed by 1e-3 misses the field
    u2 = om.propagate(tx, ty, dens * (1 + 1e-3 * np.sin(np.arange(dens.size))))
    assert np.linalg.norm(u2 - ue) / np.linalg.norm(ue) > 1e-5
