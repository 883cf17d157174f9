"""GPU: sixteen seeded random star obstacles (k = 100, ~10 nodes per
wavelength, N from ~500 to ~4,400) applied as one captured launch
(ObstacleGroup / ifgf_group_*): every obstacle's result equals its own
single application bit for bit (the same kernels), on two launches with
changed inputs and eager work in between; one obstacle checked against the
oracle."""
import numpy as np
import pytest

import ifgf2d_amd as I
import oracle as O
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu


def test_obstacle_group_launch():
    torch = pytest.importorskip("torch")
    kws = configs.random_stars(16, k=100.0, seed=3)
    maps = [I.ForwardMap(**kw) for kw in kws]
    print("obstacle sizes", [m.N for m in maps], flush=True)
    xs = [torch.zeros(m.N, dtype=torch.complex128, device="cuda:0") for m in maps]
    ys = [torch.zeros_like(x) for x in xs]
    torch.cuda.synchronize()
    group = I.ObstacleGroup(maps, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys])
    rng = np.random.default_rng(9)
    for rep in range(2):
        hx = [rng.standard_normal(m.N) + 1j * rng.standard_normal(m.N) for m in maps]
        for x, h in zip(xs, hx):
            x.copy_(torch.from_numpy(h))
        torch.cuda.synchronize()
        group.launch()
        group.synchronize()
        for m, y, h in zip(maps, ys, hx):
            assert np.array_equal(y.cpu().numpy(), m.compute_Ax_acc(h))
    kw = kws[0]
    om = O.ForwardMap(kw["curve"][0], kw["curve"][1], kw["patch_lims"], formulation=O.COMBINED, k=100.0, eta=100.0,
                      delta=maps[0].delta, p=6.0, nlevels=kw["nlevels"])
    yo = om.apply(hx[0])
    e = np.linalg.norm(ys[0].cpu().numpy() - yo) / np.linalg.norm(yo)
    print("obstacle 0 vs oracle", e, flush=True)
    assert e <= 1e-12
    group.close()


def test_group_mixes_classes_and_is_one_launch_per_stage():
    """Obstacles of different cone orders (chosen by the autotuner against an
    accuracy target), precisions, formulations and autotuned lane counts in one
    group: each class of equal (Ps, Pang) and precision runs every stage as one
    launch, and every obstacle still equals its own application bit for bit.
    The fused group is timed against the obstacles applied one after another."""
    import time
    torch = pytest.importorskip("torch")
    kws = configs.random_stars(16, k=100.0, seed=5)
    maps = []
    for j, kw in enumerate(kws):
        extra = {}
        if j % 4 == 1:
            ps, pa, err, ok = I.autotune_cone_order(kw["curve"], kw["patch_lims"], kw["k"], kw["nlevels"], 1e-5, p=6.0)
            print(f"obstacle {j}: autotuned cone order {ps}x{pa}, error {err:.1e}", flush=True)
            assert ok and err <= 1e-5
            extra = dict(Ps=ps, Pang=pa)
        if j % 4 == 2:
            extra = dict(precision="mixed")
        if j % 4 == 3:
            extra = dict(formulation="single")
        maps.append(I.ForwardMap(**kw, autotune=(j % 2 == 0), **extra))
    xs = [torch.zeros(m.N, dtype=torch.complex128, device="cuda:0") for m in maps]
    ys = [torch.zeros_like(x) for x in xs]
    torch.cuda.synchronize()
    group = I.ObstacleGroup(maps, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys])
    st = group.stats()
    print("group stats", st, flush=True)
    assert st["plans"] == 16 and st["classes"] >= 3
    maxd = max(m.nlevels for m in maps) - 1
    assert st["launches"] <= st["classes"] * (3 + maxd + 3)
    rng = np.random.default_rng(1)
    hx = [rng.standard_normal(m.N) + 1j * rng.standard_normal(m.N) for m in maps]
    for x, h in zip(xs, hx):
        x.copy_(torch.from_numpy(h))
    torch.cuda.synchronize()
    group.launch()
    group.synchronize()
    for m, y, h in zip(maps, ys, hx):
        assert np.array_equal(y.cpu().numpy(), m.compute_Ax_acc(h))
    # timing: one after another on device pointers, and the fused group
    for _ in range(3):
        for m, x, y in zip(maps, xs, ys):
            m.apply_device(x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        for m, x, y in zip(maps, xs, ys):
            m.apply_device(x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t0) / 20
    group.launch()
    group.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        group.launch()
    group.synchronize()
    t_grp = (time.perf_counter() - t0) / 20
    print(f"16 obstacles: sequential {t_seq * 1e3:.3f} ms, fused group {t_grp * 1e3:.3f} ms "
          f"({t_seq / t_grp:.1f}x)", flush=True)
    assert t_grp < t_seq
    group.close()


def test_cone_order_autotuner_picks_the_cheapest_order_within_tolerance():
    kw = configs.random_stars(1, k=100.0, seed=2)[0]
    errs = {}
    for tol in (1e-2, 1e-4, 1e-6):
        ps, pa, err, ok = I.autotune_cone_order(kw["curve"], kw["patch_lims"], kw["k"], kw["nlevels"], tol, p=6.0)
        print(f"tol {tol:.0e}: {ps}x{pa} error {err:.2e}", flush=True)
        assert ok and err <= tol
        errs[tol] = (ps * pa, err)
    assert errs[1e-2][0] <= errs[1e-4][0] <= errs[1e-6][0]
