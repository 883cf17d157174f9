"""The C ABI from a C++ caller on the GPU (tests/cpp_operator_check.cpp, built by
__graft_entry__.build()): the forward map behind a LinearOperator-style object
solved by the reference's host GMRES restatement and by the device GMRES, the
sound-soft circle's scattered field against the Mie series at 9 x 9 and 5 x 5
cones (the C1 configuration)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp_operator_check")


def test_cpp_caller_solves_the_circle_against_mie():
    if not os.path.exists(EXE):
        raise FileNotFoundError(f"{EXE} is missing: run __graft_entry__.build()")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout + out.stderr, flush=True)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok")
