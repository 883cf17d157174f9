"""GPU rectangular-polar correction (IFGF_PLAN_DEVICE_PRECOMP) against the host's.

The correction rows fold the RP precomputation (ForwardMap.tpp:287-423; 40
Fejer nodes per column, Chebyshev moments, Cheb1D::interp_1d folded in) for
every target and its own / near-singular patches. The device evaluates each
column with the same column code as the host (rp.h) but its own sin / cos /
pow and Hankel routines, so the two agree to rounding of the column
computation. For the double layer that rounding is amplified near the
singular point: (x - y).n_y is O(|x - y|^2) formed from O(|x - y|) terms, so a
last-bit change of a sine shifts it by eps / (curvature |x - y| / 2), about
1e-10 for the RP nodes closest to the singularity. The bars: single layer
rows 1e-12 relative, double layer / combined rows ROW_TOL_DLP, and the
forward map built on the device rows within the 1e-12 oracle parity bar.
"""
import numpy as np
import pytest

import ifgf2d_amd as I
import oracle as O
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu

ROW_TOL_DLP = 1e-9

CASES = [
    ("circle-cfie", I.circle(), 40, 10.0, 3, I.COMBINED),
    ("kite-dlp", I.kite(), 400, 40.0, 6, I.DOUBLE_LAYER),
    ("star-slp", I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7, I.SINGLE_LAYER),
    ("star-cfie", I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7, I.COMBINED),
]


@pytest.mark.parametrize("name,curve,npatch,k,nlev,F", CASES, ids=[c[0] for c in CASES])
def test_device_correction_matches_host(name, curve, npatch, k, nlev, F):
    lims = I.equal_parameter_lims(npatch)
    delta = I.default_delta(curve, lims)
    host = I.ForwardMap(curve, lims, k, nlev, formulation=F, delta=delta, p=6.0, host_only=True)
    dev = I.ForwardMap(curve, lims, k, nlev, formulation=F, delta=delta, p=6.0, device_precomp=True)
    ph, pth, wh = host.correction()
    pd, ptd, wd = dev.correction()
    assert np.array_equal(ph, pd) and np.array_equal(pth, ptd)
    scale = np.abs(wh).max(axis=1)
    err = (np.abs(wd - wh).max(axis=1) / scale).max()
    print(f"{name}: {wh.shape[0]} correction rows, max row difference {err:.3e}", flush=True)
    assert err <= (1e-12 if F == I.SINGLE_LAYER else ROW_TOL_DLP)
    om = O.ForwardMap(curve[0], curve[1], lims, formulation=F, k=k, eta=k, delta=delta, p=6.0, nlevels=nlev)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(dev.N) + 1j * rng.standard_normal(dev.N)
    yo = om.apply(x)
    e = np.linalg.norm(dev.compute_Ax_acc(x) - yo) / np.linalg.norm(yo)
    print(f"  forward map on device rows vs oracle {e:.3e}", flush=True)
    assert e <= 1e-12


def test_device_correction_full_size_star():
    """The C3 star (N = 256,000, k = 500): device rows against host rows."""
    kw, _ = configs.star_k500()
    host = I.ForwardMap(**kw, host_only=True)
    dev = I.ForwardMap(**kw, device_precomp=True)
    _, _, wh = host.correction()
    _, _, wd = dev.correction()
    err = (np.abs(wd - wh).max(axis=1) / np.abs(wh).max(axis=1)).max()
    print(f"star 256k: {wh.shape[0]} rows, max row difference {err:.3e}", flush=True)
    assert err <= ROW_TOL_DLP


GEO_CASES = CASES + [("random-star-" + str(i), kw["curve"], len(kw["patch_lims"]) - 1, kw["k"], kw["nlevels"],
                      I.COMBINED) for i, kw in enumerate(configs.random_stars(3, k=100.0, seed=3))]


@pytest.mark.parametrize("name,curve,npatch,k,nlev,F", GEO_CASES, ids=[c[0] for c in GEO_CASES])
def test_device_near_singular_search_matches_host(name, curve, npatch, k, nlev, F):
    """IFGF_PLAN_DEVICE_GEOMETRY: the near-singular points found on the GPU are
    the host's (the correction's target-major structure, which lists every
    (target, own or near-singular patch) pair, is identical) and the closest
    parameters agree (rows within 1e-12, the host computing the correction)."""
    lims = I.equal_parameter_lims(npatch)
    delta = I.default_delta(curve, lims)
    host = I.ForwardMap(curve, lims, k, nlev, formulation=F, delta=delta, p=6.0, host_only=True)
    dev = I.ForwardMap(curve, lims, k, nlev, formulation=F, delta=delta, p=6.0, device_geometry=True)
    ph, pth, wh = host.correction()
    pd, ptd, wd = dev.correction()
    assert np.array_equal(ph, pd) and np.array_equal(pth, ptd)
    err = (np.abs(wd - wh).max(axis=1) / np.abs(wh).max(axis=1)).max()
    print(f"{name}: {wh.shape[0]} rows, {len(pth)} correction entries, max row difference {err:.3e}")
    assert err <= 1e-12


def test_device_setup_forward_map_against_oracle():
    """All three device setup steps (sort, near-singular search, correction):
    the forward map within the oracle parity bar."""
    curve, npatch, k, nlev = I.kite(), 400, 40.0, 6
    lims = I.equal_parameter_lims(npatch)
    delta = I.default_delta(curve, lims)
    dev = I.ForwardMap(curve, lims, k, nlev, delta=delta, p=6.0, device_setup=True)
    om = O.ForwardMap(curve[0], curve[1], lims, formulation=I.COMBINED, k=k, eta=k, delta=delta, p=6.0,
                      nlevels=nlev)
    x = np.random.default_rng(8).standard_normal(dev.N) + 0j
    yo = om.apply(x)
    e = np.linalg.norm(dev.compute_Ax_acc(x) - yo) / np.linalg.norm(yo)
    print("device setup forward map vs oracle", e, dev.setup_times())
    assert e <= 1e-12


def test_device_near_singular_search_full_size_star():
    """The C3 star (N = 256,000): the same near-singular pairs as the host."""
    kw, _ = configs.star_k500()
    host = I.ForwardMap(**kw, host_only=True)
    dev = I.ForwardMap(**kw, device_setup=True)
    ph, pth, _ = host.correction()
    pd, ptd, _ = dev.correction()
    assert np.array_equal(ph, pd) and np.array_equal(pth, ptd)
    print("star setup (host-only plan)", host.setup_times())
    print("star setup (device setup)", dev.setup_times())
