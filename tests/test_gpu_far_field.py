"""Far-field electrostatic potential on the GPU through the C-ABI
(evaluate_electrostatic_potential, src/evaluator.cpp:391-479; far_field_plan,
expand_far_field and local_polynomial, src/electrostatic_potential.cpp:363-844;
target clusters, src/potential_clusters.hpp:29-105), against the oracle:

* listed points in 1.6-bohr cubes and grids in 8^3 tiles, at tolerances 1e-6
  and 1e-8, stay within the tolerance of the exact potential: the GPU's exact
  pair sums at every point and the oracle's exact sum on a sample (the
  reference's own check, tests/electrostatic_potential_test.cpp:264-348);
* the oracle's far-field values of the same targets agree with the GPU's to
  1e-11 (the same plan decisions, which pairs and source clusters are
  approximated at each target, summed in another order);
* a value depends only on its point, its target and the plan: a subset of the
  points, or a grid evaluated in ranges that cut its tiles (the serving
  loop), gives the same values bit for bit;
* tolerance 0 and the exact method give the exact path's values bit for bit.

Workloads: the 20-atom 6-31G*-shaped molecule on 100 000 points of its 1.4 x
van der Waals surface and on a grid, and the 40-atom def2-TZVP-shaped
molecule (640 AOs) on a grid over the whole molecule."""
import os

import numpy as np
import pytest

from kumo_hip import synthetic
from kumo_hip.api import PropertyKind as K
from support import oracle
from support.fchk import FchkData

pytestmark = pytest.mark.gpu

ESP = [(K.ElectrostaticPotential, 0, 0, 0)]
FAR, EXACT = 1, 2
_MOLS = {}


@pytest.fixture(scope="module")
def ctx():
    from kumo_hip.binding import Context
    c = Context(0)
    yield c
    for g, _ in _MOLS.values():
        g.close()
    _MOLS.clear()
    c.close()


def load(ctx, name):
    from kumo_hip.binding import Molecule
    if name not in _MOLS:
        if name == "pople20":
            mol = synthetic.pople_molecule()
            path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "kumo_pople20_seed3.fchk")
            mol.write_fchk(path)
        else:
            path, mol, _ = synthetic.cached_fchk(lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
        print(f"[gpu] far field: load {name}", flush=True)
        _MOLS[name] = (Molecule(ctx, fchk_path=path), oracle.Molecule(FchkData(path)))
    return _MOLS[name]


def esp(g, tol, method, **where):
    return g.evaluate(ESP, potential_tolerance=tol, esp_far_field=method, **where)[0][0][0]


def box_grid(xyz, margin, counts):
    lo = xyz.min(0) - margin
    hi = xyz.max(0) + margin
    steps = (hi - lo) / (np.asarray(counts) - 1)
    return {"origin": lo, "steps": np.diag(steps), "counts": np.asarray(counts, dtype=np.int32)}


def cube_subset(pts, ncubes, seed):
    # every point of `ncubes` of the 1.6-bohr cubes that hold points
    cells = np.floor(pts / 1.6).astype(np.int64)
    keys = np.unique(cells, axis=0)
    pick = keys[np.random.default_rng(seed).choice(len(keys), size=min(ncubes, len(keys)), replace=False)]
    mask = np.zeros(len(pts), bool)
    for k in pick:
        mask |= np.all(cells == k, axis=1)
    return np.nonzero(mask)[0]


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_vdw_surface_points(ctx, tol):
    g, o = load(ctx, "pople20")
    mol = synthetic.pople_molecule()
    pts = synthetic.vdw_shell_points(mol.atomic_numbers, mol.xyz, 100000)
    print(f"[gpu] far field: vdW surface, {len(pts)} points, tolerance {tol:g}", flush=True)
    far = esp(g, tol, FAR, points=pts)
    exact = esp(g, 0.0, EXACT, points=pts)
    err = np.abs(far - exact)
    sample = np.arange(0, len(pts), len(pts) // 300)
    ref = o.esp(pts[sample])
    err_oracle = np.abs(far[sample] - ref)
    sub = cube_subset(pts, 25, seed=7)
    far_sub = esp(g, tol, FAR, points=pts[sub])
    orc_sub = o.evaluate(ESP, points=pts[sub], potential_tolerance=tol)[0][0][0]
    print(f"[gpu] far field: max |far - exact| {err.max():.3g} (mean {err.mean():.3g}), vs oracle exact "
          f"{err_oracle.max():.3g}, vs oracle far field {np.abs(far_sub - orc_sub).max():.3g} on {len(sub)} points",
          flush=True)
    assert err.max() <= tol
    assert err_oracle.max() <= tol
    assert err.max() > 0.0  # the far field was taken
    assert np.array_equal(far_sub, far[sub])
    assert np.abs(far_sub - orc_sub).max() <= 1e-11  # the same decisions: rounding apart


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_grid_tiles(ctx, tol):
    g, o = load(ctx, "pople20")
    grid = box_grid(synthetic.pople_molecule().xyz, 3.0, (46, 41, 37))
    print(f"[gpu] far field: grid {tuple(grid['counts'])}, tolerance {tol:g}", flush=True)
    far = esp(g, tol, FAR, grid=grid)
    exact = esp(g, 0.0, EXACT, grid=grid)
    err = np.abs(far - exact)
    small = box_grid(synthetic.pople_molecule().xyz, 1.0, (19, 12, 10))
    far_small = esp(g, tol, FAR, grid=small)
    orc_small = o.evaluate(ESP, grid=small, potential_tolerance=tol)[0][0][0]
    orc_exact = o.esp(oracle.grid_points(small))
    print(f"[gpu] far field: grid max |far - exact| {err.max():.3g}; small grid vs oracle far field "
          f"{np.abs(far_small - orc_small).max():.3g}, vs oracle exact {np.abs(far_small - orc_exact).max():.3g}",
          flush=True)
    assert err.max() <= tol and err.max() > 0.0
    assert np.abs(far_small - orc_exact).max() <= tol
    assert np.abs(far_small - orc_small).max() <= 1e-11


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_tzvp640_grid(ctx, tol):
    g, o = load(ctx, "tzvp40")
    xyz = o.data.xyz
    grid = box_grid(xyz, 2.5, (40, 40, 40))
    print(f"[gpu] far field: 640-AO molecule, grid 40^3, tolerance {tol:g}", flush=True)
    far = esp(g, tol, FAR, grid=grid)
    exact = esp(g, 0.0, EXACT, grid=grid)
    err = np.abs(far - exact)
    pts = oracle.grid_points(grid)
    sample = np.arange(7, len(pts), len(pts) // 150)
    ref = o.esp(pts[sample])
    print(f"[gpu] far field: 640 AOs max |far - exact| {err.max():.3g}, vs oracle exact "
          f"{np.abs(far[sample] - ref).max():.3g}", flush=True)
    assert err.max() <= tol and err.max() > 0.0
    assert np.abs(far[sample] - ref).max() <= tol


def test_exact_methods_agree(ctx):
    g, _ = load(ctx, "pople20")
    mol = synthetic.pople_molecule()
    pts = synthetic.vdw_shell_points(mol.atomic_numbers, mol.xyz, 5000)
    exact = esp(g, 0.0, EXACT, points=pts)
    assert np.array_equal(esp(g, 0.0, FAR, points=pts), exact)
    assert np.array_equal(esp(g, 1e-6, EXACT, points=pts), exact)
    grid = box_grid(mol.xyz, 2.0, (13, 11, 9))
    assert np.array_equal(esp(g, 0.0, FAR, grid=grid), esp(g, 1e-8, EXACT, grid=grid))


def test_ranges_cut_tiles(ctx):
    # the serving loop evaluates a request in ranges of `chunk_points` points,
    # which cut the 8^3 tiles of the grid
    from kumo_hip.binding import Server
    g, _ = load(ctx, "pople20")
    grid = box_grid(synthetic.pople_molecule().xyz, 2.0, (21, 19, 17))
    whole = esp(g, 1e-7, FAR, grid=grid)
    srv = Server(ctx, workers=3, chunk_points=1024)
    try:
        res, _ = srv.submit(g, ESP, grid=grid, potential_tolerance=1e-7, esp_far_field=FAR).wait()
    finally:
        srv.close()
    assert np.array_equal(res[0][0], whole)


def test_ranges_of_listed_points(ctx):
    # the serving loop's ranges of a point list: each range groups its own
    # points into cubes, and a value depends only on its point and cube
    from kumo_hip.binding import Server
    g, _ = load(ctx, "pople20")
    mol = synthetic.pople_molecule()
    pts = synthetic.vdw_shell_points(mol.atomic_numbers, mol.xyz, 20000)
    whole = esp(g, 1e-7, FAR, points=pts)
    srv = Server(ctx, workers=3, chunk_points=1024)
    try:
        res, _ = srv.submit(g, ESP, points=pts, potential_tolerance=1e-7, esp_far_field=FAR).wait()
    finally:
        srv.close()
    assert np.array_equal(res[0][0], whole)


def test_invalid_method(ctx):
    g, _ = load(ctx, "pople20")
    with pytest.raises(Exception, match="electrostatic potential method"):
        g.evaluate(ESP, points=np.zeros((4, 3)), esp_far_field=7)


def test_high_l_pairs(ctx):
    # g and h shells: pair orders up to L = 10, expansions to order 20 (the
    # largest recursion the target kernel holds) and near-pair kernels of
    # every order
    from kumo_hip.binding import Molecule
    sm, grid = synthetic.high_l_molecule()
    g = Molecule(ctx, data=sm)
    o = oracle.Molecule(sm)
    try:
        print("[gpu] far field: high-L molecule, grid", tuple(grid["counts"]), flush=True)
        far = esp(g, 1e-6, FAR, grid=grid)
        exact = esp(g, 0.0, EXACT, grid=grid)
        orc = o.evaluate(ESP, grid=grid, potential_tolerance=1e-6)[0][0][0]
        err = np.abs(far - exact).max()
        print(f"[gpu] far field: high-L max |far - exact| {err:.3g}, vs oracle far field "
              f"{np.abs(far - orc).max():.3g}", flush=True)
        assert 0.0 < err <= 1e-6
        assert np.abs(far - orc).max() <= 1e-11
    finally:
        g.close()


def test_special_points(ctx):
    # points on the nuclei (their own charge left out), duplicates, points in
    # cubes of their own and points 1000 bohr away (every source cluster
    # translated), against the oracle's far field and the exact potential
    g, o = load(ctx, "pople20")
    xyz = o.data.xyz
    far_away = xyz.mean(0) + np.array([[1000.0, 0.0, 0.0], [0.0, -1000.0, 0.0], [577.0, 577.0, 577.0]])
    pts = np.concatenate([xyz, xyz[:3], xyz[:5] + 0.37, far_away, far_away + 0.1])
    for tol in (1e-6, 1e-3):
        far = esp(g, tol, FAR, points=pts)
        orc = o.evaluate(ESP, points=pts, potential_tolerance=tol)[0][0][0]
        exact = esp(g, 0.0, EXACT, points=pts)
        assert np.array_equal(far[len(xyz):len(xyz) + 3], far[:3])
        assert np.abs(far - orc).max() <= 1e-11 * max(1.0, np.abs(orc).max())
        assert np.abs(far - exact).max() <= tol


def test_loose_tolerance_grid(ctx):
    # 1e-3: most pairs of a tile expanded, most clusters translated
    g, o = load(ctx, "pople20")
    grid = box_grid(synthetic.pople_molecule().xyz, 1.5, (17, 15, 13))
    far = esp(g, 1e-3, FAR, grid=grid)
    exact = esp(g, 0.0, EXACT, grid=grid)
    orc = o.evaluate(ESP, grid=grid, potential_tolerance=1e-3)[0][0][0]
    assert 0.0 < np.abs(far - exact).max() <= 1e-3
    assert np.abs(far - orc).max() <= 1e-11
