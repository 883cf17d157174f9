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
