"""GPU: the mixed-precision operator (IFGF_MIXED) and FGMRES with FP64 refinement.

A mixed plan applies the leaf operator, the cone coefficients, the propagation
and the interpolation in FP32 (centre factors evaluated in FP64 and rounded);
near field, correction and accumulation stay FP64 (in 8-column blocks the
near field runs on FP32 matrix cores, its sums widened to FP64). Bars,
written per test:
the mixed operator is within MIXED_TOL of the oracle's FP64 forward map
(FP32 rounding of the far-field part only); the plan's FP64 operator
(IFGF_APPLY_FP64) is bitwise the plain plan's; FGMRES (FP64 outer Arnoldi,
inner mixed GMRES) reaches the FP64 true residual 1e-8 and the same density
as plain FP64 GMRES.
"""
import numpy as np
import pytest

import ifgf2d_amd as I
import oracle as O
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu

MIXED_TOL = 2e-6  # relative 2-norm, FP32 far field


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def density(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.mark.parametrize("curve,npatch,k,nlev,F", [
    (I.circle(), 40, 10.0, 3, I.COMBINED),
    (I.kite(), 400, 40.0, 6, I.COMBINED),
    (I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7, I.SINGLE_LAYER),
], ids=["circle-cfie", "kite-cfie", "star-slp"])
def test_mixed_operator_against_oracle(curve, npatch, k, nlev, F):
    lims = I.equal_parameter_lims(npatch)
    delta = I.default_delta(curve, lims)
    fm = I.ForwardMap(curve, lims, k, nlev, formulation=F, delta=delta, p=6.0, precision="mixed")
    f64 = I.ForwardMap(curve, lims, k, nlev, formulation=F, delta=delta, p=6.0)
    om = O.ForwardMap(curve[0], curve[1], lims, formulation=F, k=k, eta=k, delta=delta, p=6.0, nlevels=nlev)
    print(f"mixed case N={fm.N} k={k} nlev={nlev} F={F}", flush=True)
    x = density(fm.N, 21)
    yo = om.apply(x)
    ym = fm.compute_Ax_acc(x)
    e = rel(ym, yo)
    print(f"  mixed vs oracle {e:.3e}", flush=True)
    assert 0 < e <= MIXED_TOL
    # the FP64 operator of the mixed plan is the plain plan's, bit for bit
    y64 = fm.compute_Ax_acc(x, fp64=True)
    assert np.array_equal(y64, f64.compute_Ax_acc(x))
    assert rel(y64, yo) <= 1e-12
    # mixed applies are deterministic
    assert np.array_equal(ym, fm.compute_Ax_acc(x))


def test_mixed_full_size_kite():
    """64k kite at k = 200: the mixed operator's distance from the FP64 one,
    set against the IFGF approximation error itself (3e-4 against direct
    summation), so the FP32 far field costs nothing in accuracy."""
    kw, _ = configs.kite_k200()
    fm = I.ForwardMap(**kw, precision="mixed")
    x = density(fm.N, 22)
    ym = fm.compute_Ax_acc(x)
    y64 = fm.compute_Ax_acc(x, fp64=True)
    e = rel(ym, y64)
    print("kite 64k mixed vs fp64", e, flush=True)
    assert 0 < e < MIXED_TOL


BATCH_TOL = 1e-14  # relative 2-norm, FP64 column blocks against single columns


def test_fp64_column_blocks_on_mixed_plan():
    """IFGF_APPLY_FP64 with several columns on a mixed plan runs FP64 column
    blocks (8, 8, 8, 2, 1 for 27 columns; two lanes through device pointers):
    every column within BATCH_TOL of the same plan's single FP64 application
    and of a plain FP64 plan's blocks, within 1e-12 of the oracle; the device
    call equals the host call bitwise and repeats bitwise; the FP32 blocks of
    the same plan stay within MIXED_TOL of the FP64 ones."""
    torch = pytest.importorskip("torch")
    curve, npatch, k, nlev = I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7
    lims = I.equal_parameter_lims(npatch)
    delta = I.default_delta(curve, lims)
    kw = configs.scaled(curve, npatch, k, nlev, delta=delta)
    fm = I.ForwardMap(**kw, precision="mixed", nrhs_max=32)
    f64 = I.ForwardMap(**kw, nrhs_max=32)
    n, nr = fm.N, 27
    X = np.stack([density(n, 120 + s) for s in range(nr)], axis=1)
    Y = fm.compute_Ax_acc(X, fp64=True)
    e1 = max(rel(Y[:, j], fm.compute_Ax_acc(X[:, j], fp64=True)) for j in (0, 9, 17, 24, 26))
    e2 = max(rel(Y[:, j], Yp) for j, Yp in enumerate(f64.compute_Ax_acc(X).T))
    Ym = fm.compute_Ax_acc(X)
    e3 = max(rel(Ym[:, j], Y[:, j]) for j in range(nr))
    om = O.ForwardMap(curve[0], curve[1], lims, formulation=I.COMBINED, k=k, eta=k, delta=delta, p=6.0,
                      nlevels=nlev)
    e4 = max(rel(Y[:, j], om.apply(X[:, j])) for j in (3, 26))
    print(f"FP64 blocks vs single {e1:.3e}, vs plain plan {e2:.3e}, FP32 blocks vs FP64 {e3:.3e}, "
          f"vs oracle {e4:.3e}", flush=True)
    assert e1 <= BATCH_TOL and e2 <= BATCH_TOL and 0 < e3 <= MIXED_TOL and e4 <= 1e-12
    ld = n + 3
    xt = torch.zeros((nr, ld), dtype=torch.complex128, device="cuda:0")
    xt[:, :n] = torch.from_numpy(np.ascontiguousarray(X.T))
    s = torch.cuda.current_stream()
    for _ in range(2):
        yt = torch.zeros_like(xt)
        torch.cuda.synchronize()
        fm.apply_device(xt.data_ptr(), yt.data_ptr(), nrhs=nr, ld=ld, stream=s.cuda_stream, fp64=True)
        Yd = yt.cpu().numpy()
        assert np.array_equal(Yd[:, :n].T, Y) and np.all(Yd[:, n:] == 0)


def test_fgmres_circle_mie():
    """Sound-soft unit circle, k = 10: FGMRES on a mixed plan reaches the FP64
    true residual 1e-8 and the Mie series at r = 3, like FP64 GMRES."""
    from test_gpu_parity import mie
    kw, d = configs.circle_k10(Ps=9, Pang=9)
    fm = I.ForwardMap(**kw, precision="mixed")
    x, y, *_ = fm.points()
    rhs = -configs.plane_wave(x, y, kw["k"], d)
    dens, it, iit, relres, hist = fm.solve_fgmres(rhs, restart=20, tol=1e-8, inner_restart=50)
    print("FGMRES outer", it, "inner", iit, "true relres", relres, flush=True)
    assert relres < 1e-8
    th = 2 * np.pi * np.arange(64) / 64
    us = fm.propagate_unacc(3 * np.cos(th), 3 * np.sin(th), dens)
    assert rel(us, mie(3.0, th, kw["k"])) < 1e-5
    d64, it64, rr64, _ = I.ForwardMap(**kw).solve(rhs, restart=50, tol=1e-8)
    assert rel(dens, d64) < 1e-6


def test_fgmres_kite_against_fp64_gmres():
    """Kite, 400 patches, k = 40: mixed-plan FGMRES against plain FP64 GMRES on
    the same discrete system (true FP64 residual, densities), and mixed-plan
    GMRES alone stalls above the FGMRES residual (what the refinement buys)."""
    kw = configs.scaled(I.kite(), 400, 40.0, 6)
    fm = I.ForwardMap(**kw, precision="mixed")
    f64 = I.ForwardMap(**kw)
    x, y, *_ = fm.points()
    rhs = -configs.plane_wave(x, y, 40.0, (np.cos(np.pi / 4), np.sin(np.pi / 4)))
    d64, it64, rr64, _ = f64.solve(rhs, restart=50, tol=1e-8, maxit=500)
    dm, it, iit, relres, hist = fm.solve_fgmres(rhs, restart=20, tol=1e-8, inner_restart=50)
    print("fp64 GMRES", it64, rr64, "| FGMRES outer", it, "inner", iit, relres, flush=True)
    assert rr64 < 1e-8 and relres < 1e-8
    assert rel(dm, d64) < 1e-6
    # FGMRES on an FP64 plan (inner operator FP64) converges as well
    d2, it2, iit2, rr2, _ = f64.solve_fgmres(rhs, restart=20, tol=1e-8, inner_restart=50)
    assert rr2 < 1e-8 and rel(d2, d64) < 1e-6


def test_star_256k_fp64_and_mixed():
    """Star r = 1 + 0.3 cos 5t at k = 500, 25,600 patches (N = 256,000), 12
    levels, plane wave d = (1, 0): FP64 GMRES(50) and mixed-plan FGMRES both
    reach the FP64 true residual 1e-8 and agree on the density; the mixed
    operator stays within MIXED_TOL of the FP64 one; the IFGF error against
    the direct sum on 2,000 targets is at the 5x5 level."""
    kw, d = configs.star_k500()
    f64 = I.ForwardMap(**kw)
    print("star stats", f64.stats(), flush=True)
    x = density(f64.N, 23)
    y = f64.compute_Ax_acc(x)
    tg = np.arange(0, f64.N, 128)
    err = rel(y[tg], f64.compute_Ax_unacc(x, tg))
    print("IFGF 5x5 vs direct on", tg.size, "targets", err, flush=True)
    assert 1e-6 < err < 2e-3
    xg, yg, *_ = f64.points()
    rhs = -configs.plane_wave(xg, yg, kw["k"], d)
    d64, it, rr, _ = f64.solve(rhs, restart=50, tol=1e-8, maxit=2000)
    print("fp64 GMRES", it, rr, flush=True)
    assert rr < 1e-8
    del f64
    fm = I.ForwardMap(**kw, precision="mixed")
    e = rel(fm.compute_Ax_acc(x), y)
    print("mixed vs fp64 operator", e, flush=True)
    assert 0 < e < MIXED_TOL
    dm, oit, iit, rrm, _ = fm.solve_fgmres(rhs, restart=20, tol=1e-8, inner_restart=60)
    print("mixed FGMRES", oit, iit, rrm, "density diff", rel(dm, d64), flush=True)
    assert rrm < 1e-8 and rel(dm, d64) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_rhs_gmres(precision):
    """ifgf_gmres_multi: twelve incidence directions (blocks of 8 + 4 batched
    columns) solved in lockstep on the kite (400 patches, k = 40) against
    twelve single-RHS ifgf_gmres solves: each converges, the FP64 plan to the
    true residual 1e-8 with densities within 1e-7 and iteration counts within
    2 of the single solves."""
    kw = configs.scaled(I.kite(), 400, 40.0, 6)
    fm = I.ForwardMap(**kw, precision=precision, nrhs_max=8)
    x, y, *_ = fm.points()
    th = 2 * np.pi * np.arange(12) / 12
    B = np.stack([-configs.plane_wave(x, y, 40.0, (np.cos(t), np.sin(t))) for t in th], axis=1)
    X, its, rr = fm.solve_multi(B, restart=30, tol=1e-8, maxit=500)
    print(precision, "multi iterations", its.tolist(), "relres max", rr.max(), flush=True)
    for j in range(12):
        d, it, r1, _ = fm.solve(B[:, j], restart=30, tol=1e-8, maxit=500)
        assert abs(int(its[j]) - it) <= 2
        if precision == "fp64":
            assert rr[j] < 1e-8 and rel(X[:, j], d) < 1e-7
        else:  # the mixed operator alone levels off near its FP32 rounding, as single GMRES does
            assert rr[j] < 1e-6 and rel(X[:, j], d) < 1e-5


def test_multi_rhs_mixed_refinement():
    """ifgf_refine_multi on a mixed-precision plan: twelve incidence directions
    reach the FP64 true residual 1e-8 through lockstep mixed-precision GMRES
    corrections and FP64 residuals, with densities within 1e-7 of the FP64
    plan's lockstep solutions."""
    kw = configs.scaled(I.kite(), 400, 40.0, 6)
    f64 = I.ForwardMap(**kw, nrhs_max=8)
    fmx = I.ForwardMap(**kw, precision="mixed", nrhs_max=8)
    x, y, *_ = f64.points()
    th = 2 * np.pi * np.arange(12) / 12
    B = np.stack([-configs.plane_wave(x, y, 40.0, (np.cos(t), np.sin(t))) for t in th], axis=1)
    X64, its64, rr64 = f64.solve_multi(B, restart=30, tol=1e-8, maxit=500)
    X, outer, its, rr = fmx.solve_multi_refined(B, restart=30, tol=1e-8, maxit=500)
    print("refinement steps", outer, "inner iterations", its.tolist(), "relres max", rr.max(),
          "fp64 lockstep iterations", its64.tolist(), flush=True)
    assert 1 <= outer <= 4
    assert np.all(rr < 1e-8)
    for j in range(12):
        # the true residual checked against the FP64 operator here as well
        r = B[:, j] - f64.compute_Ax_acc(X[:, j])
        assert np.linalg.norm(r) / np.linalg.norm(B[:, j]) < 1e-8
        assert rel(X[:, j], X64[:, j]) < 1e-7


def test_multi_rhs_nonzero_initial_guess():
    """A nonzero initial guess (a zero one is neither copied nor applied),
    from a rough solution (the FP64 lockstep solution at 1e-3, one column left
    at zero). The lockstep FP64 GMRES stops on |w(k)| / ||r0|| < tol per
    system, so from the smaller r0 it goes further below ||b|| (every true
    residual below 1e-8 x the start's, the zero column to within one
    iteration of its zero-start count); the mixed refinement, which stops on
    the true residual against ||b||, takes fewer inner iterations on the
    warm-started columns than from zero. Every solution within 1e-7 of the
    zero-start ones."""
    kw = configs.scaled(I.kite(), 400, 40.0, 6)
    f64 = I.ForwardMap(**kw, nrhs_max=8)
    fmx = I.ForwardMap(**kw, precision="mixed", nrhs_max=8)
    x, y, *_ = f64.points()
    th = 2 * np.pi * np.arange(9) / 9
    B = np.stack([-configs.plane_wave(x, y, 40.0, (np.cos(t), np.sin(t))) for t in th], axis=1)
    X0, its0, rr0 = f64.solve_multi(B, restart=30, tol=1e-8, maxit=500)
    G, _, rrg = f64.solve_multi(B, restart=30, tol=1e-3, maxit=500)
    G[:, 4] = 0.0
    rrg[4] = 1.0
    X1, its1, rr1 = f64.solve_multi(B, restart=30, tol=1e-8, maxit=500, x0=G)
    Xz, outz, itz, rrz = fmx.solve_multi_refined(B, restart=30, tol=1e-8, maxit=500)
    Xr, outer, itr, rrr = fmx.solve_multi_refined(B, restart=30, tol=1e-8, maxit=500, x0=G)
    print("zero start", its0.tolist(), "| warm start", its1.tolist(), "| refined from zero", outz, itz.tolist(),
          "| from the warm start", outer, itr.tolist(), flush=True)
    assert np.all(rr0 < 1e-8) and np.all(rrz < 1e-8) and np.all(rrr < 1e-8)
    assert np.all(rr1 < 1e-8 * rrg * 10)
    assert abs(int(its1[4]) - int(its0[4])) <= 1
    assert outer <= outz and all(itr[j] < itz[j] for j in range(9) if j != 4)
    for j in range(9):
        assert rel(X1[:, j], X0[:, j]) < 1e-7 and rel(Xr[:, j], X0[:, j]) < 1e-7
        r = B[:, j] - f64.compute_Ax_acc(Xr[:, j])
        assert np.linalg.norm(r) / np.linalg.norm(B[:, j]) < 1e-8


def test_star_256k_64_directions():
    """C4 at its stated scale: 64 incidence directions 2 pi j / 64 on the
    256,000-point star at k = 500, solved together by the FP64 lockstep GMRES
    (ifgf_gmres_multi) and by the mixed-precision refinement
    (ifgf_refine_multi): every true FP64 residual below 1e-8, four columns
    within 1e-7 of single FP64 ifgf_gmres solves, and the scattered field on
    r = 3 (64 points, propagate_unacc) of every mixed-refined column within
    1e-6 of the FP64 solution's."""
    kw, _ = configs.star_k500(nrhs_max=64)
    f64 = I.ForwardMap(**kw, device_setup=True)
    x, y, *_ = f64.points()
    th = 2 * np.pi * np.arange(64) / 64
    B = np.stack([-configs.plane_wave(x, y, kw["k"], (np.cos(t), np.sin(t))) for t in th], axis=1)
    X64, its, rr = f64.solve_multi(B, restart=50, tol=1e-8, maxit=2000)
    print(f"FP64 lockstep: iterations {its.min()}..{its.max()}, max true relres {rr.max():.2e}", flush=True)
    assert np.all(rr < 1e-8)
    for j in (0, 13, 32, 57):
        d, it, r1, _ = f64.solve(B[:, j], restart=50, tol=1e-8, maxit=2000)
        print(f"  column {j}: single {it} iterations vs lockstep {its[j]}, density diff {rel(X64[:, j], d):.1e}",
              flush=True)
        assert rel(X64[:, j], d) < 1e-7
    tt = 2 * np.pi * np.arange(64) / 64
    tx, ty = 3.0 * np.cos(tt), 3.0 * np.sin(tt)
    far64 = [f64.propagate_unacc(tx, ty, X64[:, j]) for j in range(64)]
    del f64
    fmx = I.ForwardMap(**kw, precision="mixed", device_setup=True)
    Xm, outer, itm, rrm = fmx.solve_multi_refined(B, restart=50, tol=1e-8, maxit=2000)
    print(f"mixed refinement: {outer} steps, inner {itm.min()}..{itm.max()}, max true relres {rrm.max():.2e}",
          flush=True)
    assert np.all(rrm < 1e-8)
    efar = max(rel(fmx.propagate_unacc(tx, ty, Xm[:, j]), far64[j]) for j in range(64))
    eden = max(rel(Xm[:, j], X64[:, j]) for j in range(64))
    print(f"  max far-field diff vs FP64 {efar:.1e}, max density diff {eden:.1e}", flush=True)
    assert efar < 1e-6 and eden < 1e-6
