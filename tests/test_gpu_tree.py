"""Tree tables built on the GPU (IFGF_PLAN_DEVICE_TREE) against the host's.

The grouping step (every target's box per level, Level::Point2Morton; the
cousin lists, Level::GetCousins, and neighbourhoods, GetNeighbours; the
interpolation entry counts) and the cone step of the tree setup
(BoxTree::InitializeRelevantConeSegments,
BoxTree.h:553-583, with the protoboxes of GetRelevantConeSegmentsDueToParents,
:299-428) decides which cones every level keeps, in which order, which cone
every interpolation entry reads and which child cone every parent node is
propagated from. The device builder runs the host's geometry functions and
numbering order, so the plans must agree exactly: the same work counts, the
same cones per level, and applications that are bitwise equal (every table
above feeds the stored operators, so a single differing entry changes bits).
"""
import numpy as np
import pytest

import ifgf2d_amd as I
import oracle as O
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu

CASES = [
    ("circle-nlev3", I.circle(), 40, 10.0, 3),
    ("circle-nlev5", I.circle(), 200, 40.0, 5),
    ("kite-nlev6", I.kite(), 400, 40.0, 6),
    ("star-nlev7", I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7),
] + [("random-star-" + str(i), kw["curve"], len(kw["patch_lims"]) - 1, kw["k"], kw["nlevels"])
     for i, kw in enumerate(configs.random_stars(3, k=100.0, seed=5))]


def _same_plan(host, dev, nlev):
    assert host.stats() == dev.stats()
    for level in range(2, nlev):
        bh, qh = host.level_cones(level)
        bd, qd = dev.level_cones(level)
        assert np.array_equal(bh, bd) and np.array_equal(qh, qd), f"level {level}"
    rng = np.random.default_rng(11)
    x = rng.standard_normal(host.N) + 1j * rng.standard_normal(host.N)
    yh, yd = host.compute_Ax_acc(x), dev.compute_Ax_acc(x)
    assert np.array_equal(yh, yd)
    return x, yd


@pytest.mark.parametrize("name,curve,npatch,k,nlev", CASES, ids=[c[0] for c in CASES])
def test_device_cones_match_host(name, curve, npatch, k, nlev):
    lims = I.equal_parameter_lims(npatch)
    kw = dict(curve=curve, patch_lims=lims, k=k, nlevels=nlev, p=6.0)
    host = I.ForwardMap(**kw)
    dev = I.ForwardMap(**kw, device_tree=True)
    _same_plan(host, dev, nlev)
    print(f"{name}: {dev.stats()['cones']} cones, {dev.stats()['interp']} interpolation entries, "
          f"{dev.stats()['prop']} propagations; tree {dev.setup_times()['tree']:.4f} s "
          f"(host {host.setup_times()['tree']:.4f} s)", flush=True)


def test_device_cones_forward_map_against_oracle():
    """A plan whose cones were built on the GPU within the oracle parity bar."""
    curve, npatch, k, nlev = I.kite(), 400, 40.0, 6
    lims = I.equal_parameter_lims(npatch)
    delta = I.default_delta(curve, lims)
    dev = I.ForwardMap(curve, lims, k, nlev, delta=delta, p=6.0, device_tree=True)
    om = O.ForwardMap(curve[0], curve[1], lims, formulation=I.COMBINED, k=k, eta=k, delta=delta, p=6.0,
                      nlevels=nlev)
    x = np.random.default_rng(12).standard_normal(dev.N) + 0j
    yo = om.apply(x)
    e = np.linalg.norm(dev.compute_Ax_acc(x) - yo) / np.linalg.norm(yo)
    print("device cones forward map vs oracle", e)
    assert e <= 1e-12


def test_device_cones_field_matches_host():
    """The field tree of a plan with device cones (separate targets: levels may
    keep no cone) is built on the GPU too and evaluates bitwise like the host's."""
    curve, npatch, k, nlev = I.kite(), 400, 40.0, 6
    lims = I.equal_parameter_lims(npatch)
    kw = dict(curve=curve, patch_lims=lims, k=k, nlevels=nlev, p=6.0)
    host = I.ForwardMap(**kw)
    dev = I.ForwardMap(**kw, device_tree=True)
    g = np.linspace(-2.5, 2.5, 61)
    tx, ty = [a.ravel() for a in np.meshgrid(g, g)]
    rho = np.random.default_rng(13).standard_normal(host.N) + 1j
    for n in (5, 7):
        fh = host.field(tx, ty, nlevels=n)
        fd = dev.field(tx, ty, nlevels=n)
        assert fh.stats() == fd.stats()
        assert np.array_equal(fh.evaluate(rho), fd.evaluate(rho))
        fh.close()
        fd.close()


def test_device_cones_full_size_star():
    """The C3 star (N = 256,000, k = 500, 12 levels) with every setup step on
    the GPU: the same plan as the host-built one."""
    kw, _ = configs.star_k500()
    host = I.ForwardMap(**kw)
    dev = I.ForwardMap(**kw, device_setup=True)
    assert host.stats() == dev.stats()
    for level in range(2, kw["nlevels"]):
        bh, qh = host.level_cones(level)
        bd, qd = dev.level_cones(level)
        assert np.array_equal(bh, bd) and np.array_equal(qh, qd), f"level {level}"
    x = np.random.default_rng(14).standard_normal(host.N) + 1j
    yh, yd = host.compute_Ax_acc(x), dev.compute_Ax_acc(x)
    # the device correction rows differ from the host's in the last bits
    e = np.linalg.norm(yd - yh) / np.linalg.norm(yh)
    print("star 256k: host setup", host.setup_times(), flush=True)
    print("star 256k: device setup", dev.setup_times(), "difference", e, flush=True)
    assert e <= 1e-12
    tree_only = I.ForwardMap(**kw, device_tree=True)
    assert np.array_equal(tree_only.compute_Ax_acc(x), yh)
    print("star 256k: device cones only", tree_only.setup_times(), flush=True)
