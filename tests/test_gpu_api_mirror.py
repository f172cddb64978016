"""The C++ mirror of the reference Evaluator API (chem-kumo_amd/include/chem/
kumo/hip.hpp) compiled into tests/cpp/api_mirror_test and run on the GPU:
Molecule::from_fchk, Grid::around, Evaluator::evaluate on a grid and at
points, backend_used(), and the exception types of the reference."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "cpp", "api_mirror_test")


@pytest.mark.gpu
def test_cpp_api_mirror():
    fchk = os.path.join(ROOT, "tests", "golden", "fchk", "h2o_b3lyp_def2tzvp.fchk.gz")
    assert os.path.exists(BINARY), "build it with make -C chem-kumo_amd"
    r = subprocess.run([BINARY, fchk], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
