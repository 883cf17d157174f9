import os
import sys

import pytest

# torch bundles its own libamdhip64 (same SONAME as ROCm's). Importing it
# before libifgf2d.so is loaded makes both share one HIP runtime, so device
# pointers can be passed between them.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def hankel_golden():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "hankel_mpmath.npy"))
