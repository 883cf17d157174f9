"""The plan's host setup is independent of the OpenMP team it gets: the same
star plan (800 patches, k = 60, 7 levels, CFIE 5 x 5, stored operators on
the device, device setup steps off so the host's parallel loops -- the
counting sort of the push entries among them -- build every table) is built
in fresh processes with the default team, OMP_NUM_THREADS=1,
OMP_THREAD_LIMIT=2 and OMP_DYNAMIC=true, and its application to a seeded
density is bitwise the same in all four."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = """
import sys
import numpy as np
sys.path.insert(0, {root!r})
import ifgf2d_amd as I
from ifgf2d_amd import configs
kw = configs.scaled(I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7)
fm = I.ForwardMap(**kw)
y = fm.compute_Ax_acc(configs.seeded_density(fm.N, seed=3))
np.save({out!r}, y)
print("built", fm.N, flush=True)
"""


def test_plan_build_independent_of_openmp_team(tmp_path):
    outs = []
    for extra in ({}, {"OMP_NUM_THREADS": "1"}, {"OMP_THREAD_LIMIT": "2"}, {"OMP_DYNAMIC": "true"}):
        out = str(tmp_path / f"y{len(outs)}.npy")
        env = dict(os.environ, **extra)
        print(f"building with {extra or 'the default team'}", flush=True)
        r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT, out=out)], env=env, capture_output=True,
                           text=True, timeout=120)
        print(r.stdout + r.stderr[-2000:], flush=True)
        assert r.returncode == 0
        outs.append(np.load(out))
    for extra, y in zip(("OMP_NUM_THREADS=1", "OMP_THREAD_LIMIT=2", "OMP_DYNAMIC=true"), outs[1:]):
        assert np.array_equal(y, outs[0]), extra
