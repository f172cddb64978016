"""The C++ mirror of the reference Evaluator API (chem-kumo_amd/include/chem/
kumo/hip.hpp) compiled into tests/cpp/api_mirror_test and run on the GPU:
Molecule::from_fchk, Grid::around, Evaluator::evaluate on a grid and at
points, backend_used(), Field::write_cube / write_orbital_cube (the case of
tests/atom_charge_cube_test.cpp), the molecular surface functions of
include/chem/kumo/surface.hpp, and the exception types of the reference."""
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
    # The ESP surface through the mirror: the oracle's mesh of the same
    # isosurface (rho = 0.001, 0.25 bohr) and its area.
    line = [x for x in r.stdout.splitlines() if x.startswith("esp_surface ")][0].split()
    vals = dict(kv.split("=") for kv in line[1:])
    v, t = oracle.density_isosurface(oracle.Molecule(d), 1e-3, 0.25)
    assert (int(vals["vertices"]), int(vals["triangles"])) == (len(v), len(t))
    assert float(vals["area"]) == pytest.approx(oracle.mesh_area(v, t), rel=1e-9)
