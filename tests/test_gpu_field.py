"""GPU parity of the IFGF field evaluator at separate targets (ifgf_field_*).

Bar: the FP64 field matches the oracle's separate-target BoxTree (the
reference's BoxTree::Solve with the targets in place of the points, itself
pinned against BoxTree::Solve and against propagate_unacc's direct sum) to
<= 1e-12 relative (2-norm) on the same seeded inputs; the mixed-precision
field matches the FP64 one to 2e-6. At the 64k-point kite the IFGF field is
checked against the GPU direct sum (propagate_unacc) on a target subset.
"""
import numpy as np
import pytest

import ifgf2d_amd as I
import oracle as O
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-12
MIXED_TOL = 2e-6


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def density(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def targets(which, x, y, seed=5):
    rng = np.random.default_rng(seed)
    if which == "ring":
        t = rng.uniform(0, 2 * np.pi, 500)
        return 1.5 * np.cos(t), 1.5 * np.sin(t)
    if which == "grid":
        g = np.linspace(-2.0, 2.0, 33)
        return tuple(a.ravel() for a in np.meshgrid(g, g))
    if which == "points":
        return x.copy(), y.copy()
    if which == "mixed-set":  # on-surface, near-surface, interior and far targets together
        t = rng.uniform(0, 2 * np.pi, 300)
        r = np.concatenate([np.ones(100), 1 + rng.uniform(-0.05, 0.05, 100), rng.uniform(0, 6, 100)])
        return np.concatenate([r * np.cos(t), x[::7]]), np.concatenate([r * np.sin(t), y[::7]])
    raise ValueError(which)


CASES = [
    # target set, formulation, Ps, Pang
    ("ring", I.COMBINED, 5, 5),
    ("grid", I.COMBINED, 5, 5),
    ("points", I.COMBINED, 5, 5),
    ("mixed-set", I.COMBINED, 5, 5),
    ("grid", I.SINGLE_LAYER, 5, 5),
    ("grid", I.DOUBLE_LAYER, 5, 5),
    ("ring", I.COMBINED, 9, 9),
    ("mixed-set", I.SINGLE_LAYER, 7, 7),
]


@pytest.mark.parametrize("which,F,Ps,Pang", CASES, ids=[f"{c[0]}-{c[1]}-{c[2]}x{c[3]}" for c in CASES])
def test_field_matches_oracle(which, F, Ps, Pang):
    fm = I.ForwardMap(I.circle(), I.equal_parameter_lims(80), 10.0, 4, formulation=F)
    x, y, nx, ny, w = fm.points()
    tx, ty = targets(which, x, y)
    u = density(fm.N, 1)
    f = fm.field(tx, ty, Ps=Ps, Pang=Pang)
    out = f.evaluate(u)
    ref, so = O.field_solve(x, y, nx, ny, u * w, tx, ty, f.nlevels, 10.0, F, fm.eta, Ps=Ps, Pang=Pang)
    st = f.stats()
    for key in ("near", "leaf", "prop", "interp", "cones"):
        assert st[key] == so[key], key
    e = rel(out, ref)
    print(which, F, Ps, Pang, "levels", f.nlevels, "rel err vs oracle", e)
    assert e < FIELD_TOL


def test_field_kite_matches_oracle_and_direct():
    fm = I.ForwardMap(I.kite(), I.equal_parameter_lims(400), 40.0, 6)
    x, y, nx, ny, w = fm.points()
    g = np.linspace(-2.2, 2.2, 64)
    tx, ty = (a.ravel() for a in np.meshgrid(g, g))
    u = density(fm.N, 2)
    for P in (5, 9):
        f = fm.field(tx, ty, Ps=P, Pang=P)
        out = f.evaluate(u)
        ref, _ = O.field_solve(x, y, nx, ny, u * w, tx, ty, f.nlevels, 40.0, I.COMBINED, fm.eta, Ps=P, Pang=P)
        direct = fm.propagate_unacc(tx, ty, u)
        e_or, e_dir = rel(out, ref), rel(out, direct)
        print("kite field", P, "levels", f.nlevels, "vs oracle", e_or, "vs direct", e_dir)
        assert e_or < FIELD_TOL
        assert e_dir < (1e-2 if P == 5 else 1e-5)


def test_field_mixed_precision_and_device_pointers():
    torch = pytest.importorskip("torch")
    fm = I.ForwardMap(I.kite(), I.equal_parameter_lims(400), 40.0, 6)
    x, y = fm.points()[:2]
    tx, ty = targets("mixed-set", x, y)
    u = density(fm.N, 3)
    f64 = fm.field(tx, ty)
    fmx = fm.field(tx, ty, precision="mixed")
    a = f64.evaluate(u)
    b = fmx.evaluate(u)
    c = fmx.evaluate(u, fp64=True)
    print("mixed field rel diff", rel(b, a), "fp64 operator of the mixed field", rel(c, a))
    assert rel(b, a) < MIXED_TOL
    assert rel(c, a) < 1e-14
    # bitwise determinism and the device-pointer path
    assert np.array_equal(a, f64.evaluate(u))
    dev = torch.device("cuda")
    ud = torch.from_numpy(u).to(dev)
    od = torch.empty(f64.M, dtype=torch.complex128, device=dev)
    s = torch.cuda.current_stream()
    f64.apply_device(ud.data_ptr(), od.data_ptr(), stream=s.cuda_stream)
    s.synchronize()
    assert np.array_equal(od.cpu().numpy(), a)


def test_field_single_and_far_targets():
    fm = I.ForwardMap(I.circle(), I.equal_parameter_lims(40), 10.0, 3)
    x, y, nx, ny, w = fm.points()
    u = density(fm.N, 4)
    for tx, ty in ((np.array([2.0]), np.array([0.5])), (np.array([40.0, -30.0, 0.0]), np.array([0.0, 25.0, 0.0]))):
        f = fm.field(tx, ty)
        out = f.evaluate(u)
        ref, _ = O.field_solve(x, y, nx, ny, u * w, tx, ty, f.nlevels, 10.0, I.COMBINED, fm.eta)
        print("levels", f.nlevels, rel(out, ref))
        assert rel(out, ref) < FIELD_TOL
        assert rel(out, fm.propagate_unacc(tx, ty, u)) < 1e-2


def test_field_kite_64k_against_direct_subset():
    """The 64k-point kite at k = 200 with a 256 x 256 target grid over the
    scatterer: IFGF field (5x5 and 9x9) against the GPU direct sum on 2,000
    targets."""
    kw, _ = configs.kite_k200()
    fm = I.ForwardMap(**kw)
    g = np.linspace(-2.0, 2.0, 256)
    tx, ty = (a.ravel() for a in np.meshgrid(g, g))
    u = density(fm.N, 6)
    sub = np.random.default_rng(7).choice(tx.size, 2000, replace=False)
    direct = fm.propagate_unacc(tx[sub], ty[sub], u)
    errs = {}
    for P in (5, 9):
        f = fm.field(tx, ty, Ps=P, Pang=P)
        out = f.evaluate(u)
        errs[P] = rel(out[sub], direct)
        print("64k kite field", P, f.stats(), "rel err vs direct", errs[P])
        assert np.array_equal(out, f.evaluate(u))
    assert errs[5] < 1e-2 and errs[9] < 1e-5 and errs[9] < errs[5] / 10


def test_field_of_autotuned_plan():
    """A tuned plan's field is tuned too (task sizes, lanes, near field on a
    second stream); it agrees with the untuned field to rounding."""
    curve, lims = I.kite(), I.equal_parameter_lims(400)
    fa = I.ForwardMap(curve, lims, 40.0, 6)
    fb = I.ForwardMap(curve, lims, 40.0, 6, autotune=True)
    x, y = fa.points()[:2]
    tx, ty = targets("grid", x, y)
    u = density(fa.N, 13)
    a = fa.field(tx, ty).evaluate(u)
    b = fb.field(tx, ty).evaluate(u)
    print("tuned field vs default", rel(b, a))
    assert rel(b, a) < 1e-13
