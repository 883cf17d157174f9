"""The GPU forward map at the full C2 and C3 sizes against the oracle, and the
whole solve pipeline against a closed form at those sizes.

* The 64,000-point kite (k = 200) and the 256,000-point star (k = 500), CFIE
  5 x 5, plans built the way the benchmark builds them (every setup step on
  the GPU, autotuned): the GPU application of the seeded density within 1e-12
  of the oracle's (OpenMP over the host threads; its results do not depend on
  the thread count).
* Interior point source x0: the radiating exterior solution of the Dirichlet
  problem with data g = H0(k |x - x0|) on the curve is u = H0(k |x - x0|)
  itself, so the CFIE density that GMRES finds must reproduce u at points on
  r = 3 (ForwardMap::solve then propagate_unacc, ForwardMap.tpp:658-727).
  9 x 9 cones keep the IFGF error below the 1e-5 bar; the FP64 GMRES and the
  mixed-precision FGMRES (FP32 cone work, FP64 refinement) are both checked.
"""
import numpy as np
import pytest
import scipy.special as sp

import ifgf2d_amd as I
import oracle as O
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle_of(kw, delta):
    c = kw["curve"]
    return O.ForwardMap(c[0], c[1], kw["patch_lims"], formulation=O.COMBINED, k=kw["k"], eta=kw["k"], delta=delta,
                        p=kw["p"], nlevels=kw["nlevels"])


@pytest.mark.parametrize("which", ["kite_k200", "star_k500"])
def test_full_size_operator_matches_oracle(which):
    kw, _ = getattr(configs, which)()
    fm = I.ForwardMap(**kw, autotune=True, device_setup=True)
    print(f"{which}: N={fm.N} threads for the oracle: {O.set_threads(0)}", flush=True)
    x = configs.seeded_density(fm.N)
    y = fm.compute_Ax_acc(x)
    om = _oracle_of(kw, fm.delta)
    yo = om.apply(x)
    e = rel(y, yo)
    print(f"{which}: GPU vs oracle {e:.2e}", flush=True)
    assert e < 1e-12
    st, so = fm.stats(), om.stats()
    assert all(st[c] == so[c] for c in ("near", "leaf", "prop", "interp", "cones"))


def _point_source(x, y, x0, k):
    return sp.hankel1(0, k * np.hypot(x - x0[0], y - x0[1]))


@pytest.mark.parametrize("which,x0", [("kite_k200", (-0.2, 0.3)), ("star_k500", (0.15, -0.25))])
def test_full_size_point_source_closed_form(which, x0):
    kw, _ = getattr(configs, which)()
    k = kw["k"]
    th = 2 * np.pi * np.arange(64) / 64
    tx, ty = 3.0 * np.cos(th), 3.0 * np.sin(th)
    ue = _point_source(tx, ty, x0, k)
    fm = I.ForwardMap(**kw, Ps=9, Pang=9, device_setup=True)
    x, y, *_ = fm.points()
    g = _point_source(x, y, x0, k)
    dens, it, rr, _ = fm.solve(g, restart=50, tol=1e-10, maxit=600)
    u = fm.propagate_unacc(tx, ty, dens)
    e64 = rel(u, ue)
    print(f"{which} 9x9 FP64 GMRES: {it} iterations, true relres {rr:.1e}, field error {e64:.2e}", flush=True)
    assert rr < 1e-9 and e64 < 1e-5
    del fm
    fmx = I.ForwardMap(**kw, Ps=9, Pang=9, precision="mixed", device_setup=True)
    dm, oit, iit, rrm, _ = fmx.solve_fgmres(g, restart=20, tol=1e-10, inner_restart=60)
    um = fmx.propagate_unacc(tx, ty, dm)
    em = rel(um, ue)
    print(f"{which} 9x9 mixed FGMRES: {oit} outer + {iit} inner, true relres {rrm:.1e}, field error {em:.2e}, "
          f"vs FP64 field {rel(um, u):.1e}", flush=True)
    assert rrm < 1e-9 and em < 1e-5 and rel(um, u) < 1e-8
