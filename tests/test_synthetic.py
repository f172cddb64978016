"""The synthetic FCHK generator: its analytic overlap agrees with the
oracle's (src/basis_evaluator.cpp:210-278 restated), the contracted AOs are
normalised, the MOs are S-orthonormal, the FCHK text round-trips through the
reader, and the occupied MOs reproduce P (AO-MO eligible)."""
import os

import numpy as np

from kumo_hip import synthetic
from support import oracle
from support.fchk import FchkData


def test_config0_molecule(tmp_path):
    mol, grid = synthetic.water_like(seed=0)
    assert mol.nbasis == 19
    path = mol.write_fchk(str(tmp_path / "w.fchk"))
    d = FchkData(path)
    assert d.nbasis == 19 and d.nalpha == 5 and d.nbeta == 5
    om = oracle.Molecule(d)
    S = om.basis.overlap()
    assert np.allclose(S, mol.overlap, rtol=1e-12, atol=1e-13)
    assert np.allclose(np.diag(S), 1.0, atol=1e-12)
    C = d.mo_matrix(0)
    assert np.allclose(C.T @ S @ C, np.eye(19), atol=1e-9)
    assert om.eligible
    n = int(np.prod(grid["counts"]))
    assert 250_000 < n < 500_000


def test_tzvp_shapes_small(tmp_path):
    mol, grid = synthetic.batch_molecule(seed=3, natoms=5, family="tzvp", edge=40)
    d = FchkData(mol.write_fchk(str(tmp_path / "t.fchk")))
    om = oracle.Molecule(d)
    assert np.allclose(om.basis.overlap(), mol.overlap, rtol=1e-11, atol=1e-12)
    assert om.eligible
    assert list(grid["counts"]) == [40, 40, 40]
