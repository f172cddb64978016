"""Parity of the HIP path (through the C-ABI) with the CPU oracle and with the
reference's external known answers, on the reference's own fixture
molecules. FP64 contractions: values within 1e-11 relative of the oracle
(different summation order only); screening masks, point order and AO
indexing bit-exact."""
import sys

import numpy as np
import pytest

from kumo_hip.api import Backend, DensityComponent as DC, Precision, PropertyKind as K
from support import oracle
from support.fchk import FchkData
from support.reference_cases import (DENSITY_CUBES, FCHK, MO_CUBES, MULTIWFN, check_density_cube, check_mo_cube,
                                     check_multiwfn, load_sample)

pytestmark = pytest.mark.gpu

_OR = {}
_GPU = {}


@pytest.fixture(scope="module")
def ctx():
    from kumo_hip.binding import Context
    c = Context(0)
    yield c
    for m in _GPU.values():
        m.close()
    _GPU.clear()
    c.close()


def omol(name):
    if name not in _OR:
        _OR[name] = oracle.Molecule(FchkData(FCHK[name]))
    return _OR[name]


def gmol(ctx, name):
    from kumo_hip.binding import Molecule
    if name not in _GPU:
        print(f"[gpu] load {name}", flush=True)
        _GPU[name] = Molecule(ctx, fchk_path=FCHK[name])
    return _GPU[name]


def rel_err(a, b, floor=1e-300):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


@pytest.mark.parametrize("name", ["o2", "acrolein", "water"])
def test_prepared_blocks_bit_exact(ctx, name):
    g = gmol(ctx, name).blocks()
    o = omol(name).basis.blocks()
    for a, b in zip(g, o):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("name", ["water", "acrolein"])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_ao_values_match_oracle(ctx, name, order):
    rng = np.random.default_rng(order)
    pts = rng.normal(size=(257, 3)) * 2.0
    g = gmol(ctx, name).ao(pts, order)
    o = omol(name).basis.ao(pts, order)
    scale = np.abs(o).max(axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(g - o) / scale) < 1e-13


@pytest.mark.parametrize("name", ["o2", "acrolein", "water"])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("block", [128, 37])
def test_masks_bit_exact(ctx, name, order, block):
    d = load_sample("o2_density" if name == "o2" else "acrolein_density")
    grid = d["grid"] if name != "water" else {"origin": np.array([-4.0, -3.5, -3.0]), "steps": np.eye(3) * 0.21,
                                              "counts": np.array([30, 28, 26])}
    gm = gmol(ctx, name).masks(block, order, grid=grid)
    om = omol(name).masks(block, order, grid=grid)
    assert gm.shape == om.shape
    assert np.array_equal(gm, om)


@pytest.mark.parametrize("block", [64, 300])
@pytest.mark.parametrize("order", [0, 2])
@pytest.mark.parametrize("threshold", [1e-12, 1e-6])
def test_masks_bit_exact_points(ctx, block, order, threshold):
    # listed points, blocks below and above 128 points, and a coarse threshold
    rng = np.random.default_rng(5 + block)
    pts = rng.normal(size=(1500, 3)) * 3.0
    gm = gmol(ctx, "acrolein").masks(block, order, points=pts, threshold=threshold)
    om = omol("acrolein").masks(block, order, points=pts, threshold=threshold)
    assert np.array_equal(gm, om)


def test_boys_matches_oracle(ctx):
    rng = np.random.default_rng(1)
    T = np.concatenate([[0.0, 1e-10, 0.025, 39.99, 40.0, 55.5], rng.uniform(0, 60, 2000)])
    g = ctx.boys(T, 16)
    o = np.array([oracle.boys(t, 16) for t in T])
    assert rel_err(g, o) < 1e-13


@pytest.mark.parametrize("name", DENSITY_CUBES)
def test_density_against_cubegen_and_oracle(ctx, name):
    d = load_sample(name)
    comp = int(d["component"])
    mol = str(d["molecule"])
    for backend in (Backend.AO_P, Backend.AUTO):
        print(f"[gpu] density {name} backend={int(backend)}", flush=True)
        g, used = gmol(ctx, mol).evaluate([(K.Density, comp, 0, 0)], grid=d["grid"], backend=backend)
        o, oused = omol(mol).evaluate([(K.Density, comp, 0, 0)], grid=d["grid"], backend=backend)
        assert used == oused
        scale = np.abs(o[0][0]).max()
        assert np.max(np.abs(g[0][0] - o[0][0])) <= 1e-12 * scale
        check_density_cube(name, g[0][0])


@pytest.mark.parametrize("name", MO_CUBES)
def test_orbitals_against_cubegen_and_oracle(ctx, name):
    d = load_sample(name)
    mol = str(d["molecule"])
    prop = [(K.Orbital, 0, int(d["channel"]), int(d["orbital"]))]
    g, _ = gmol(ctx, mol).evaluate(prop, grid=d["grid"])
    o, _ = omol(mol).evaluate(prop, grid=d["grid"])
    assert np.max(np.abs(g[0][0] - o[0][0])) <= 1e-12 * np.abs(o[0][0]).max()
    check_mo_cube(name, g[0][0])


@pytest.mark.parametrize("case", MULTIWFN, ids=[c[0] for c in MULTIWFN])
def test_derived_fields_against_multiwfn_and_oracle(ctx, case):
    mol = case[1]
    print(f"[gpu] multiwfn {case[0]}", flush=True)
    gm = gmol(ctx, mol)
    got = {}

    def ev(grid, props):
        g, _ = gm.evaluate(props, grid=grid)
        got["g"], got["grid"], got["props"] = g, grid, props
        return g
    check_multiwfn(case, ev, omol(mol).data.xyz)
    o, _ = omol(mol).evaluate(got["props"], grid=got["grid"])
    for a, b, p in zip(got["g"], o, got["props"]):
        scale = np.abs(b).max() + 1e-300
        if p[0] == K.SignLambda2Density:
            # sign(lambda2) is ill-posed where lambda2 ~ 0: count, do not demand
            mism = np.sum(np.sign(a) != np.sign(b))
            assert mism <= 2, mism
            continue
        assert np.max(np.abs(a - b)) <= 1e-9 * scale, (case[0], p)


@pytest.mark.parametrize("name", ["o2", "water"])
def test_point_path_and_backends(ctx, name):
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(1000, 3)) * 1.7
    props = [(K.Density, c, 0, 0) for c in range(4)] + [(K.Gradient, 0, 0, 0), (K.Hessian, 1, 0, 0),
                                                        (K.Laplacian, 3, 0, 0), (K.Tau, 2, 0, 0), (K.Elf, 0, 0, 0),
                                                        (K.Lol, 1, 0, 0), (K.ReducedGradient, 0, 0, 0),
                                                        (K.TauComponents, 0, 0, 0),
                                                        (K.HamiltonianKineticEnergyComponents, 0, 0, 0),
                                                        (K.Orbital, 0, 0, 3), (K.OrbitalGradient, 0, 1, 2),
                                                        (K.OrbitalHessian, 0, 0, 1)]
    for backend in (Backend.AO_P, Backend.AO_MO):
        g, used = gmol(ctx, name).evaluate(props, points=pts, backend=backend)
        o, oused = omol(name).evaluate(props, points=pts, backend=backend)
        assert used == oused == backend
        for a, b, p in zip(g, o, props):
            assert np.max(np.abs(a - b)) <= 1e-10 * (np.abs(b).max() + 1e-300), (backend, p)


def test_block_size_and_batching_invariance(ctx):
    rng = np.random.default_rng(9)
    pts = rng.normal(size=(3001, 3))
    props = [(K.Density, 0, 0, 0), (K.Elf, 0, 0, 0), (K.Hessian, 0, 0, 0)]
    m = gmol(ctx, "acrolein")
    base, _ = m.evaluate(props, points=pts, block_size=128)
    o, _ = omol("acrolein").evaluate(props, points=pts, block_size=128)
    for bs in (1, 37, 128, 3006):
        g, _ = m.evaluate(props, points=pts, block_size=bs)
        ob, _ = omol("acrolein").evaluate(props, points=pts, block_size=bs)
        for a, b in zip(g, ob):
            assert np.max(np.abs(a - b)) <= 1e-10 * (np.abs(b).max() + 1e-300), bs
    # the same block evaluated alone and inside a larger batch is bit-identical
    part, _ = m.evaluate(props, points=pts[:1280], block_size=128)
    for a, b in zip(part, base):
        assert np.array_equal(a, b[:, :1280])


@pytest.mark.parametrize("precision,rho_tol", [(Precision.FP32, 1e-6), (Precision.BF16X6, 1e-6),
                                               (Precision.BF16X3, 2e-3)])
def test_reduced_precision_errors(ctx, precision, rho_tol):
    """Mixed-precision contractions against the FP64 oracle; the bound is on
    the relative density error where rho >= 1e-6 and the ELF absolute error
    where rho >= 1e-4 (tests/derived_properties_test.cpp:188-200). With the
    FP64 rescue of points whose density or ELF the contraction leaves
    uncertain, FP32 and BF16X6 keep the density within 1e-6 and the ELF within
    1e-5 at every point, tails included (BF16X3 has no density rescue)."""
    d = load_sample("acrolein_density")
    grid = {"origin": d["origin"] + 2.0, "steps": np.eye(3) * 0.2, "counts": np.array([40, 38, 30])}
    props = [(K.Density, 0, 0, 0), (K.Elf, 0, 0, 0), (K.Laplacian, 0, 0, 0)]
    g, _ = gmol(ctx, "acrolein").evaluate(props, grid=grid, precision=precision)
    o, _ = omol("acrolein").evaluate(props, grid=grid)
    rho = o[0][0]
    sig = rho >= 1e-6
    rel = np.max(np.abs(g[0][0] - rho)[sig] / rho[sig])
    elf_err = np.max(np.abs(g[1][0] - o[1][0])[rho >= 1e-4])
    elf_all = np.max(np.abs(g[1][0] - o[1][0]))
    # the Laplacian relative to its largest |value| (no rescue is aimed at it)
    lap = np.max(np.abs(g[2][0] - o[2][0])) / np.max(np.abs(o[2][0]))
    print(f"[gpu] precision {int(precision)}: max rel rho {rel:.3g}, ELF abs {elf_err:.3g} (all points {elf_all:.3g}), "
          f"Laplacian {lap:.3g}", flush=True)
    assert rel < rho_tol
    assert elf_err < 50 * rho_tol
    assert elf_all < 1e-5
    assert lap < rho_tol


def test_reduced_precision_spin_elf(ctx):
    """BF16X6 ELF of each spin channel and of the total on the UHF O2 molecule:
    the rescue tests each channel's own error scales (the alpha and beta
    intervals, and both together for the total), so every point stays within
    1e-5 of the oracle, and the channel density within 1e-6 where rho >= 1e-6."""
    d = load_sample("o2_density")
    grid = {"origin": d["origin"] + 1.0, "steps": np.eye(3) * 0.2, "counts": np.array([36, 34, 40])}
    props = [(K.Elf, DC.Alpha, 0, 0), (K.Elf, DC.Beta, 0, 0), (K.Elf, 0, 0, 0), (K.Density, DC.Alpha, 0, 0)]
    g, _ = gmol(ctx, "o2").evaluate(props, grid=grid, precision=Precision.BF16X6)
    o, _ = omol("o2").evaluate(props, grid=grid)
    errs = [float(np.max(np.abs(g[i][0] - o[i][0]))) for i in range(3)]
    rho = o[3][0]
    sig = rho >= 1e-6
    rel = float(np.max(np.abs(g[3][0] - rho)[sig] / rho[sig]))
    print(f"[gpu] O2 BF16X6 ELF alpha/beta/total {errs}, alpha density rel {rel:.3g}", flush=True)
    assert max(errs) < 1e-5
    assert rel < 1e-6


def test_esp_matches_oracle(ctx):
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(400, 3)) * 2.5
    g, _ = gmol(ctx, "water").evaluate([(K.ElectrostaticPotential, 0, 0, 0)], points=pts)
    o = omol("water").esp(pts)
    assert np.max(np.abs(g[0][0] - o)) <= 1e-12 * np.abs(o).max()


def test_errors_map_to_exceptions(ctx):
    m = gmol(ctx, "water")
    pts = np.zeros((2, 3))
    with pytest.raises(ValueError):
        m.evaluate([(K.Elf, DC.Difference, 0, 0)], points=pts)
    with pytest.raises(ValueError):
        m.evaluate([(K.Orbital, 0, 0, 0)], points=pts)
    with pytest.raises(IndexError):
        m.evaluate([(K.Orbital, 0, 0, 100000)], points=pts)
    with pytest.raises(ValueError):
        m.evaluate([(K.Density, 0, 0, 0)], points=pts, block_size=0)
    with pytest.raises(ValueError):
        m.evaluate([(K.Density, 0, 0, 0)], points=np.array([[np.nan, 0, 0]]))
    with pytest.raises(ValueError, match="the local electron affinity is defined only for the total density"):
        m.evaluate([(K.LocalElectronAffinity, DC.Alpha, 0, 0)], points=pts)
    with pytest.raises(RuntimeError):
        m.evaluate([(K.Density, 0, 0, 0)], points=pts, progress=lambda d, t: True)


@pytest.mark.parametrize("precision", [Precision.FP32, Precision.BF16X3, Precision.BF16X6])
def test_tile_shapes_give_identical_results(ctx, precision):
    """The tile kernel's AO->MO tile (16 or 32 points per workgroup, or the
    autotuner's choice) changes only the work partition: results are
    bit-identical. The split-bf16 precisions take the wave-uniform kernel
    (tile 64) by default except for full Hessians: its products are the tile
    kernel's, its Phi arithmetic its own, so it agrees with the tile kernel to
    the contraction precision (relative to each output's largest |value|:
    BF16X6 2e-6, BF16X3 2e-3) and with itself run to run."""
    rng = np.random.default_rng(12)
    pts = rng.normal(size=(300 * 128 + 45, 3)) * 2.0  # above the tuner's batch threshold
    m = gmol(ctx, "acrolein")
    wave = precision != Precision.FP32
    tol = 2e-6 if precision == Precision.BF16X6 else 2e-3
    for props in ([(K.Density, 0, 0, 0)], [(K.Gradient, 1, 0, 0)], [(K.Laplacian, 0, 0, 0), (K.Elf, 0, 0, 0)],
                  [(K.Hessian, 0, 0, 0), (K.Density, 2, 0, 0)]):
        hessian = props[0][0] == K.Hessian
        tiles = (16, 32) if wave and not hessian else (16, 32, 0)
        got = [m.evaluate(props, points=pts, precision=precision, backend=Backend.AO_MO, tile_points=t)[0]
               for t in tiles]
        for other in got[1:]:
            for a, b in zip(got[0], other):
                assert np.array_equal(a, b), (precision, props)
        if wave and not hessian:
            w64 = [m.evaluate(props, points=pts, precision=precision, backend=Backend.AO_MO, tile_points=t)[0]
                   for t in (64, 0, 64)]
            for a, b in zip(w64[0], w64[1]):
                assert np.array_equal(a, b), (precision, props)
            for a, b in zip(w64[0], w64[2]):
                assert np.array_equal(a, b), (precision, props)
            for prop, a, b in zip(props, w64[0], got[0]):
                if prop[0] == K.Elf:
                    assert np.max(np.abs(a - b)) < 1e-5, (precision, props)
                else:
                    assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b)), (precision, props)
