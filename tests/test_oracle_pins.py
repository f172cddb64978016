"""Pins of the CPU oracle before it is trusted as the checker of the HIP
path. Each check is chosen so that a plausible slip (a dropped term, a wrong
sign or index, a transposed operand) fails it:

* Boys function: bit-identical to the reference's own src/boys_function.cpp
  (compiled into oracle/_ref), plus F_m(0) = 1/(2m+1), the erf closed form of
  F_0 and quadrature (tests/boys_function_test.cpp).
* Pure transform and primitive normalisation: analytic overlap of a single
  shell is the identity; contracted AOs of the FCHK bases have unit norm.
* AO derivatives: central finite differences of values and gradients.
* Densities, spin densities and MO amplitudes: Gaussian 09 cubegen cubes
  (tests/cube_reference_test.cpp, tests/orbital_reference_test.cpp).
* Derived fields and the ESP: Multiwfn 2026.9.1 cubes
  (tests/multiwfn_reference_test.cpp).
* Invariants: total = alpha + beta, Laplacian = trace of the Hessian, AO-P and
  AO-MO agree, ELF in [0, 1].
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from kumo_hip.api import Backend, DensityComponent as DC, PropertyKind as K
from support import oracle
from support.fchk import FchkData
from support.reference_cases import (DENSITY_CUBES, FCHK, MO_CUBES, MULTIWFN, ORBITAL_ENERGY, ORBITAL_WEIGHTED,
                                     check_density_cube, check_mo_cube, check_multiwfn, check_orbital_weighted,
                                     load_sample)

_MOLS = {}


def molecule(name):
    if name not in _MOLS:
        _MOLS[name] = oracle.Molecule(FchkData(FCHK[name]))
    return _MOLS[name]


# ---------------------------------------------------------------- Boys
def test_boys_matches_reference_bit_for_bit():
    ref = oracle.ref_boys_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes
    rng = np.random.default_rng(7)
    Ts = np.concatenate([[0.0, 1e-12, 0.025, 0.05, 39.975, 39.99999, 40.0, 40.5, 1e3],
                         rng.uniform(0, 45, 400), 10 ** rng.uniform(-8, 3, 100)])
    for T in Ts:
        for m in (0, 1, 4, 8, 16, 32):
            a = oracle.boys(T, m)
            b = np.zeros(m + 1)
            assert ref.ref_boys(float(T), m, b.ctypes.data_as(ctypes.c_void_p)) == 0
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), (T, m)


def test_boys_closed_forms_and_quadrature():
    for m in range(33):
        assert oracle.boys(0.0, m)[m] == pytest.approx(1.0 / (2 * m + 1), rel=1e-14)
    for T in (1e-6, 0.3, 1.7, 12.3, 39.9, 40.1, 77.0):
        f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        assert oracle.boys(T, 0)[0] == pytest.approx(f0, rel=2e-14)
    for T in (0.4, 3.3, 25.0, 41.0):
        vals = oracle.boys(T, 12)
        for m in (1, 5, 12):
            q = integrate.quad(lambda t: t ** (2 * m) * math.exp(-T * t * t), 0, 1, epsabs=0, epsrel=1e-13)[0]
            assert vals[m] == pytest.approx(q, rel=1e-12)


# ---------------------------------------------------------------- basis
def test_pure_transform_orthonormal_single_shell():
    class One:
        pass
    for l in (2, 3, 4, 5):
        d = One()
        d.shell_types = np.array([-l], np.int32)
        d.nprim = np.array([1], np.int32)
        d.shell_atom = np.array([1], np.int32)
        d.shell_xyz = np.array([[0.3, -0.2, 0.1]])
        d.exponents = np.array([0.77])
        d.coefficients = np.array([1.0])
        d.p_coefficients = None
        b = oracle.Basis(d)
        S = b.overlap()
        assert b.nao == 2 * l + 1
        assert np.allclose(S, np.eye(2 * l + 1), atol=1e-12), l
        # Cartesian shells: diagonal of the normalised primitives is 1
        d.shell_types = np.array([l], np.int32)
        Sc = oracle.Basis(d).overlap()
        assert np.allclose(np.diag(Sc), 1.0, atol=1e-12)


@pytest.mark.parametrize("name", ["o2", "acrolein", "water"])
def test_contracted_aos_normalised(name):
    S = molecule(name).basis.overlap()
    assert np.allclose(np.diag(S), 1.0, atol=5e-6)
    assert np.allclose(S, S.T, atol=1e-14)


@pytest.mark.parametrize("name", ["water", "acrolein"])
def test_ao_derivatives_finite_differences(name):
    b = molecule(name).basis
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(6, 3)) * 1.5
    h = 1e-5
    d2 = b.ao(pts, 2)  # nao x 10n
    n = pts.shape[0]
    for axis in range(3):
        e = np.zeros(3)
        e[axis] = h
        vp = b.ao(pts + e, 1)
        vm = b.ao(pts - e, 1)
        fd_grad = (vp[:, :n] - vm[:, :n]) / (2 * h)
        assert np.allclose(d2[:, (1 + axis) * n:(2 + axis) * n], fd_grad, rtol=3e-3, atol=3e-6)
        # second derivatives: d/d(axis) of gradient component c
        for c in range(3):
            col = [[0, 1, 2], [1, 3, 4], [2, 4, 5]][axis][c]
            fd = (vp[:, (1 + c) * n:(2 + c) * n] - vm[:, (1 + c) * n:(2 + c) * n]) / (2 * h)
            assert np.allclose(d2[:, (4 + col) * n:(5 + col) * n], fd, rtol=3e-2, atol=3e-5)
    # value panel of the derivative kernel equals the value kernel to rounding
    v0 = b.ao(pts, 0)
    assert np.allclose(d2[:, :n], v0, rtol=1e-12, atol=1e-300)


def test_shell_cutoff_bounds_values():
    """Just inside a block's cutoff radius its AOs and their first and second
    derivatives are at most 1e-14 (include/chem/kumo/basis_evaluator.hpp:62-66);
    just outside they are exact zeros."""
    d = FchkData(FCHK["water"])
    b = molecule("water").basis
    l, off, cnt, r2 = b.blocks()
    centers = []
    for s, t in enumerate(d.shell_types):
        centers += [d.shell_xyz[s]] * (2 if t == -1 else 1)
    rng = np.random.default_rng(5)
    dirs = rng.normal(size=(16, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    for blk in range(len(l)):
        r = math.sqrt(r2[blk])
        rows = slice(off[blk], off[blk] + cnt[blk])
        inside = b.ao(centers[blk] + dirs * r * (1 - 1e-7), 2)[rows]
        outside = b.ao(centers[blk] + dirs * r * (1 + 1e-7), 2)[rows]
        assert np.abs(inside).max() <= 1e-14
        assert np.abs(inside).max() > 0.0
        assert np.all(outside == 0.0)


# ---------------------------------------------------------------- cubegen
@pytest.mark.parametrize("name", DENSITY_CUBES)
def test_oracle_density_against_cubegen(name):
    d = load_sample(name)
    mol = molecule(str(d["molecule"]))
    comp = int(d["component"])
    outs, used = mol.evaluate([(K.Density, comp, 0, 0)], grid=d["grid"], backend=Backend.AO_P)
    assert used == Backend.AO_P
    check_density_cube(name, outs[0][0])


@pytest.mark.parametrize("name", ["o2_density", "o2_spin"])
def test_oracle_density_ao_mo_against_cubegen(name):
    d = load_sample(name)
    mol = molecule(str(d["molecule"]))
    outs, used = mol.evaluate([(K.Density, int(d["component"]), 0, 0)], grid=d["grid"])
    assert used == Backend.AO_MO
    check_density_cube(name, outs[0][0])


@pytest.mark.parametrize("name", MO_CUBES)
def test_oracle_orbitals_against_cubegen(name):
    d = load_sample(name)
    mol = molecule(str(d["molecule"]))
    outs, _ = mol.evaluate([(K.Orbital, 0, int(d["channel"]), int(d["orbital"]))], grid=d["grid"])
    check_mo_cube(name, outs[0][0])


# ---------------------------------------------------------------- Multiwfn
@pytest.mark.parametrize("case", MULTIWFN, ids=[c[0] for c in MULTIWFN])
def test_oracle_against_multiwfn(case):
    mol = molecule(case[1])

    def ev(grid, props):
        return mol.evaluate(props, grid=grid)[0]
    check_multiwfn(case, ev, mol.data.xyz)


# ------------------------------------------------- orbital-energy fields
@pytest.mark.parametrize("case", ORBITAL_ENERGY, ids=[c[0] for c in ORBITAL_ENERGY])
def test_oracle_orbital_energy_fields_against_multiwfn(case):
    mol = molecule(case[1])

    def ev(grid, props):
        return mol.evaluate(props, grid=grid)[0]
    check_multiwfn(case, ev, mol.data.xyz)


@pytest.mark.parametrize("case", ORBITAL_WEIGHTED, ids=[c[0] for c in ORBITAL_WEIGHTED])
def test_oracle_orbital_weighted_fukui_against_multiwfn(case):
    mol = molecule(case[1])

    def ev(grid, props):
        return mol.evaluate(props, grid=grid)[0]
    check_orbital_weighted(case, ev, np.asarray(mol.data.e_alpha), mol.data.nalpha)


def test_oracle_orbital_energy_identities_and_errors():
    """chi = (ALIE + LEA) / 2 and eta = (ALIE - LEA) / 2 with ALIE > 0 where
    rho > 0; f0 = (f+ + f-) / 2, dual = f+ - f-; f+ and f- integrate to about
    one electron (each set of weights is normalized); the errors of
    src/orbital_energy_fields.hpp:117-135 for an open shell and a nonpositive
    delta (src/evaluator.cpp:532-533)."""
    mol = molecule("acrolein")
    grid = {"origin": [-6.0, -6.0, -4.0], "steps": np.eye(3) * 0.4, "counts": [31, 31, 21]}
    kinds = [K.AverageLocalIonizationEnergy, K.LocalElectronAffinity, K.LocalElectronegativity, K.LocalHardness,
             K.OrbitalWeightedFukuiPlus, K.OrbitalWeightedFukuiMinus, K.OrbitalWeightedFukuiZero,
             K.OrbitalWeightedDualDescriptor, K.Density]
    o, _ = mol.evaluate([(k, 0, 0, 0) for k in kinds], grid=grid)
    alie, lea, chi, eta, fp, fm, f0, dual, rho = (x[0] for x in o)
    assert np.allclose(chi, 0.5 * (alie + lea), rtol=1e-14, atol=1e-16)
    assert np.allclose(eta, 0.5 * (alie - lea), rtol=1e-14, atol=1e-16)
    assert np.all(alie[rho > 1e-30] > 0)
    assert np.array_equal(f0, 0.5 * (fp + fm)) and np.array_equal(dual, fp - fm)
    dv = 0.4 ** 3
    assert abs(fp.sum() * dv - 1.0) < 0.05 and abs(fm.sum() * dv - 1.0) < 0.05
    with pytest.raises(ValueError, match="defined only for a restricted closed-shell calculation"):
        molecule("o2").evaluate([(K.OrbitalWeightedFukuiPlus, 0, 0, 0)], points=np.zeros((1, 3)))
    with pytest.raises(ValueError, match="fukui_delta must be finite and positive"):
        mol.evaluate([(K.AverageLocalIonizationEnergy, 0, 0, 0)], points=np.zeros((1, 3)), fukui_delta=0.0)
    with pytest.raises(ValueError, match="the local electron affinity is defined only for the total density"):
        mol.evaluate([(K.LocalElectronAffinity, 1, 0, 0)], points=np.zeros((1, 3)))


# ---------------------------------------------------------------- invariants
def test_invariants_total_alpha_beta_trace():
    mol = molecule("o2")
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(300, 3)) * 1.5
    props = [(K.Density, c, 0, 0) for c in range(4)] + [(K.Hessian, 0, 0, 0), (K.Laplacian, 0, 0, 0),
                                                        (K.Elf, 0, 0, 0), (K.Tau, 0, 0, 0), (K.TauComponents, 0, 0, 0)]
    for backend in (Backend.AO_P, Backend.AO_MO):
        o, used = mol.evaluate(props, points=pts, backend=backend)
        assert used == backend
        t, a, b, s = (o[i][0] for i in range(4))
        assert np.allclose(t, a + b, rtol=1e-9, atol=1e-14)
        assert np.allclose(s, a - b, rtol=1e-9, atol=1e-12)
        h = o[4]
        assert np.allclose(o[5][0], h[0] + h[3] + h[5], rtol=1e-12, atol=1e-14)
        assert np.all((o[6][0] >= 0) & (o[6][0] <= 1))
        assert np.allclose(o[7][0], o[8].sum(axis=0), rtol=1e-12, atol=1e-15)
    o_p, _ = mol.evaluate(props, points=pts, backend=Backend.AO_P)
    o_m, _ = mol.evaluate(props, points=pts, backend=Backend.AO_MO)
    for i in range(len(props)):
        assert np.allclose(o_p[i], o_m[i], rtol=2e-5, atol=1e-9), props[i]


def test_screening_changes_little_and_masks_depend_on_order():
    mol = molecule("acrolein")
    d = load_sample("acrolein_density")
    pts = oracle.grid_points(d["grid"], 0, 5000)
    exact, _ = mol.evaluate([(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0)], points=pts, threshold=0.0)
    scr, _ = mol.evaluate([(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0)], points=pts)
    assert np.allclose(scr[0], exact[0], rtol=0, atol=1e-10)
    m0 = mol.masks(128, 0, points=pts)
    m2 = mol.masks(128, 2, points=pts)
    assert np.all(m2[m0 == 1] == 1)  # a value above threshold keeps the row at every order
    # derivative panels keep AOs whose values alone are below the threshold
    assert np.count_nonzero((m2 == 1) & (m0 == 0)) > 0
    # the definition (src/ao_screening.hpp:35-50) by brute force: the largest
    # |entry| of the block's value (order 0) or value + gradient + Hessian
    # panels (order 2) above the threshold
    for b in range(m0.shape[0]):
        p = pts[b * 128:(b + 1) * 128]
        for order, m in ((0, m0), (2, m2)):
            largest = np.abs(mol.basis.ao(p, order)).max(axis=1)
            assert np.array_equal((largest > 1e-12).astype(np.uint8), m[b]), (b, order)
    loose, _ = mol.evaluate([(K.Density, 0, 0, 0)], points=pts, threshold=1e-3)
    assert not np.allclose(loose[0], exact[0], rtol=0, atol=1e-10)


def test_block_size_invariance():
    mol = molecule("water")
    rng = np.random.default_rng(2)
    pts = rng.normal(size=(301, 3))
    props = [(K.Density, 0, 0, 0), (K.Elf, 0, 0, 0), (K.Hessian, 0, 0, 0)]
    base, _ = mol.evaluate(props, points=pts, block_size=128)
    for bs in (1, 37, 306):
        o, _ = mol.evaluate(props, points=pts, block_size=bs)
        for i in range(len(props)):
            assert np.allclose(o[i], base[i], rtol=1e-11, atol=1e-14)


def test_errors_match_reference_contract():
    mol = molecule("water")
    pts = np.zeros((2, 3))
    with pytest.raises(ValueError):
        mol.evaluate([(K.Elf, DC.Difference, 0, 0)], points=pts)
    with pytest.raises(ValueError):
        mol.evaluate([(K.ReducedGradient, DC.Alpha, 0, 0)], points=pts)
    with pytest.raises(ValueError):
        mol.evaluate([(K.Orbital, 0, 0, 0)], points=pts)
    with pytest.raises(IndexError):
        mol.evaluate([(K.Orbital, 0, 0, 10_000)], points=pts)
    with pytest.raises(ValueError):
        mol.evaluate([(K.Density, 0, 0, 0)], points=pts, block_size=0)


def test_elf_tail_is_undetermined_in_fp64():
    """ELF = 1 / (1 + (D / D0)^2) with D = tau - |grad rho|^2 / 8 rho per spin
    channel (src/evaluator.cpp:127-140). Far from the nuclei one AO dominates
    every occupied MO of a channel, D vanishes in exact arithmetic and what
    FP64 computes is a rounding of tau (~1e-16 tau), while D0 / tau ~ rho^(2/3)
    falls below 1e-16 near rho ~ 1e-20: there ELF is not determined by FP64
    arithmetic. Shown with the oracle's two FP64 formulations of the
    reference algorithm (AO-MO: sums over MOs; AO-P: P times the AO panels)
    on the 640-AO UHF molecule of the 160^3 benchmark cube, at the blocks of
    its lowest densities: they disagree by more than 0.5 below rho = 1e-20 and
    agree within 1e-6 where rho >= 1e-14, the floor above which the GPU paths
    are held to the oracle's ELF."""
    from kumo_hip import synthetic
    path, _, grid = synthetic.cached_fchk(lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
    mol = oracle.Molecule(FchkData(path))
    # 300 of the screening blocks that hold points with rho < 1e-14
    r, _ = mol.evaluate([(K.Density, 0, 0, 0)], grid=grid)
    blocks = np.unique(np.nonzero(r[0][0] < 1e-14)[0] // 128)
    blocks = np.sort(np.random.default_rng(0).choice(blocks, size=300, replace=False))
    idx = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
    pts = oracle.grid_points(grid)[idx]
    props = [(K.Density, 0, 0, 0), (K.Elf, 0, 0, 0)]
    mo, _ = mol.evaluate(props, points=pts, backend=Backend.AO_MO)
    ap, _ = mol.evaluate(props, points=pts, backend=Backend.AO_P)
    rho = mo[0][0]
    d = np.abs(mo[1][0] - ap[1][0])
    assert np.allclose(rho, ap[0][0], rtol=1e-12, atol=0)
    assert d[rho < 1e-20].max() > 0.5
    assert d[rho >= 1e-14].max() < 1e-6
