"""GPU Morton box assignment and stable point sort (ifgf_morton_sort), bit-exact.

BoxTree::SortBox (BoxTree.h:73-106) assigns every point its leaf box
floor((x - min) / box size) in the bounding box ComputeBB enlarges by 1e-10
(BoxTree.h:179-202), interleaves the box indices (x in the low bit,
Level.h:172-237) and stable-sorts the points by code. Two independent
checkers: the host tree's permutation (the plan's own std::stable_sort) and a
numpy restatement below (IEEE FP64 operations, np.argsort(kind="stable")).
The bar is exact equality of integers.
"""
import numpy as np
import pytest

import ifgf2d_amd as I
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu


def spread(v):
    v = v.astype(np.uint64)
    for sh, m in ((16, 0xFFFF0000FFFF), (8, 0xFF00FF00FF00FF), (4, 0xF0F0F0F0F0F0F0F),
                  (2, 0x3333333333333333), (1, 0x5555555555555555)):
        v = (v | (v << np.uint64(sh))) & np.uint64(m)
    return v


def numpy_sort(x, y, nlevels):
    D = nlevels - 1
    mn0, mn1 = x.min() - 1e-10, y.min() - 1e-10
    mx0, mx1 = x.max() + 1e-10, y.max() + 1e-10
    bs = max(mx0 - mn0, mx1 - mn1) / float(1 << D)
    px = np.floor((x - mn0) / bs).astype(np.int64)  # all >= 0: trunc == floor
    py = np.floor((y - mn1) / bs).astype(np.int64)
    code = (spread(px) | (spread(py) << np.uint64(1))).astype(np.int64)
    return np.argsort(code, kind="stable"), code


GEOMS = [
    ("circle-400", I.circle(), 40, 3),
    ("kite-4000", I.kite(), 400, 6),
    ("star-8000", I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 7),
    ("kite-64000", I.kite(), 6400, 10),
    ("star-256000", I.star(a=[0, 0, 0, 0, 0, 0.3]), 25600, 12),
]


@pytest.mark.parametrize("name,curve,npatch,nlev", GEOMS, ids=[g[0] for g in GEOMS])
def test_device_sort_matches_host_and_numpy(name, curve, npatch, nlev):
    lims = I.equal_parameter_lims(npatch)
    fm = I.ForwardMap(curve, lims, 10.0, nlev, host_only=True)
    x, y, *_ = fm.points()
    print(f"{name}: N={x.size} nlevels={nlev}", flush=True)
    s, c = I.morton_sort(x, y, nlev)
    sn, cn = numpy_sort(x, y, nlev)
    assert np.array_equal(c, cn)
    assert np.array_equal(s, sn)
    assert np.array_equal(s, fm.sorting())


def test_device_sort_ties_keep_input_order():
    """Many points per leaf box (and exact duplicates): ties keep input order."""
    rng = np.random.default_rng(5)
    n = 100_003
    x = np.round(rng.uniform(-2, 3, n), 2)
    y = np.round(rng.uniform(-1, 1, n), 2)
    x[::7] = x[0]
    y[::7] = y[0]
    for nlev in (3, 9, 16):
        s, c = I.morton_sort(x, y, nlev)
        sn, cn = numpy_sort(x, y, nlev)
        assert np.array_equal(c, cn) and np.array_equal(s, sn), nlev


def test_plan_with_device_sort_is_identical():
    """A plan whose tree uses the device permutation applies bit-identically."""
    kw = configs.scaled(I.kite(), 400, 40.0, 6)
    a = I.ForwardMap(**kw)
    b = I.ForwardMap(**kw, device_sort=True)
    assert np.array_equal(a.sorting(), b.sorting())
    rng = np.random.default_rng(6)
    x = rng.standard_normal(a.N) + 1j * rng.standard_normal(a.N)
    assert np.array_equal(a.compute_Ax_acc(x), b.compute_Ax_acc(x))
