"""The C++ mirror of the reference Evaluator API (chem-kumo_amd/include/chem/
kumo/hip.hpp) compiled into tests/cpp/api_mirror_test and run on the GPU:
Molecule::from_fchk, Grid::around, Evaluator::evaluate on a grid and at
points, backend_used(), Field::write_cube / write_orbital_cube (the case of
tests/atom_charge_cube_test.cpp), and the exception types of the reference."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "cpp", "api_mirror_test")


@pytest.mark.gpu
def test_cpp_api_mirror(tmp_path):
    fchk = os.path.join(ROOT, "tests", "golden", "fchk", "h2o_b3lyp_def2tzvp.fchk.gz")
    assert os.path.exists(BINARY), "build it with make -C chem-kumo_amd"
    r = subprocess.run([BINARY, fchk, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    # The integral printed by the C++ program equals the oracle's on the same
    # Grid::around(atoms, 3.0, 0.3) grid.
    import numpy as np
    from kumo_hip import synthetic
    from support import oracle
    from support.fchk import FchkData
    d = FchkData(fchk)
    grid = synthetic.grid_around(d.xyz, 3.0, 0.3)
    line = [x for x in r.stdout.splitlines() if x.startswith("grid ")][0].split()
    assert [int(v) for v in line[1:4]] == list(grid["counts"])
    assert np.array_equal(np.array([float(v) for v in line[5:8]]), grid["origin"])
    got = float([x for x in r.stdout.splitlines() if x.startswith("density_sum=")][0].split()[0].split("=")[1])
    ref, _ = oracle.Molecule(d).evaluate([(0, 0, 0, 0)], grid=grid)
    assert abs(got - ref[0][0].sum()) <= 1e-12 * ref[0][0].sum()
