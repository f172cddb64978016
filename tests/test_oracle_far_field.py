"""Pins of the oracle's far-field electrostatic potential
(src/electrostatic_potential.cpp:363-844; target clusters
src/potential_clusters.hpp:29-105, src/evaluator.cpp:391-479), after the
reference's own checks (tests/electrostatic_potential_test.cpp):

* tolerance 0 plans nothing and equals the exact pointwise sum bit for bit
  (:225-262);
* cluster values stay within the tolerance of the exact potential around
  centres in, near and far from the molecule, and some pairs are expanded
  (:305-368); at 1000 bohr every source cluster goes through its multipole
  translation and no pair is summed exactly;
* the source-cluster moments carry the electron count, dipole and traceless
  quadrupole of the analytic AO moment integrals (a separate code path,
  tests/support/moments.py), and the exact potential at 1000 bohr follows
  that multipole expansion (:98-138);
* the evaluator's grid tiles and point cubes at the default tolerance 1e-8
  stay within 1e-8 of the exact values, independently of the block size,
  and a grid point on a nucleus gets the potential at the nucleus (:264-303);
* invalid use is rejected (:350-368).

A slip in the bounds would show as an error above the tolerance (or no
expansion at all), a slip in the moments or translations as a wrong multipole
limit or cluster value, a slip in the tile or cube geometry as a point outside
its planned radius (rejected) or an error above the tolerance."""
import numpy as np
import pytest

from kumo_hip import synthetic
from kumo_hip.api import PropertyKind as K
from support import moments, oracle
from support.fchk import FchkData
from support.reference_cases import FCHK

_MOLS = {}


def molecule(name):
    if name not in _MOLS:
        if name in FCHK:
            _MOLS[name] = oracle.Molecule(FchkData(FCHK[name]))
        elif name == "pople20":
            _MOLS[name] = oracle.Molecule(synthetic.pople_molecule())
        else:
            _MOLS[name] = oracle.Molecule(synthetic.high_l_molecule()[0])
    return _MOLS[name]


NAMES = ["water", "o2", "acrolein", "high_l"]
ESP = [(K.ElectrostaticPotential, 0, 0, 0)]


def _offsets(radius):
    # the centre, then the 26 directions of a cube alternately on the sphere and at 0.6 of it
    out, q = [], 0
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                v = np.array([i, j, k], float)
                scale = 0.0 if not v.any() else (1.0 if q % 2 else 0.6)
                out.append(v / (np.linalg.norm(v) if v.any() else 1.0) * radius * scale)
                q += 1
    return np.array(out)


def _centres(xyz):
    return [xyz[0] + np.array([0.3, 0.2, -0.4]), xyz[-1] + np.array([-2.0, 1.0, 1.5]), np.array([9.0, -7.0, 4.0])]


@pytest.mark.parametrize("name", NAMES)
def test_tolerance_zero_is_exact(name):
    m = molecule(name)
    pot = m.potential()
    plan = pot.plan(1.2, 0.0)
    assert plan.finite_levels == 0 and plan.nclusters == 0
    for c in _centres(m.data.xyz):
        pts = _offsets(1.2) + c
        v, stats = pot.cluster(pts, c, plan)
        assert stats == (0, 0, plan.npairs)
        assert np.array_equal(v, m.esp(pts))
    grid = {"origin": np.array([-3.1, -2.9, -3.3]), "steps": np.array([[0.61, 0.05, 0], [0, 0.61, 0], [0, 0, 0.61]]),
            "counts": np.array([9, 8, 11])}
    got = m.evaluate(ESP, grid=grid, potential_tolerance=0.0)[0][0][0]
    assert np.array_equal(got, m.esp(oracle.grid_points(grid)))


@pytest.mark.parametrize("tol", [1e-4, 1e-6, 1e-8])
@pytest.mark.parametrize("name", NAMES)
def test_clusters_within_tolerance(name, tol):
    m = molecule(name)
    pot = m.potential()
    radius = 1.2
    plan = pot.plan(radius, tol)
    assert plan.finite_levels > 0 and plan.nclusters > 0
    assert plan.pair_tolerance == 0.5 * tol
    worst, expanded = 0.0, 0
    for c in _centres(m.data.xyz):
        pts = _offsets(radius) + c
        v, (translated, exp, exact) = pot.cluster(pts, c, plan)
        expanded += translated + exp
        worst = max(worst, float(np.max(np.abs(v - m.esp(pts)))))
    assert expanded > 0
    assert worst <= tol, worst
    # far away every source cluster is one multipole translation
    c = np.array([600.0, 700.0, -300.0]) + m.data.xyz.mean(0)
    pts = _offsets(radius) + c
    v, (translated, exp, exact) = pot.cluster(pts, c, plan)
    assert translated == plan.nclusters and exp == 0 and exact == 0
    assert np.max(np.abs(v - m.esp(pts))) <= tol


def _traceless(s):
    S = np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
    return 3.0 * S - np.eye(3) * np.trace(S)


@pytest.mark.parametrize("name", NAMES)
def test_cluster_moments_match_analytic_ao_moments(name):
    m = molecule(name)
    charge, dipole, second = moments.electron_moments(m.data)
    plan = m.potential().plan(1.0, 1e-8)
    q, d, s = 0.0, np.zeros(3), np.zeros(6)
    pairs = 0
    pos = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    for k in range(plan.nclusters):
        info = plan.cluster_info(k)
        c, mom = info["centre"], info["moments"]
        pairs += info["npairs"]
        assert info["radius"] <= 2.0 * np.sqrt(3.0) and info["min_distance"] > info["radius"]
        # moments Q_nu = sum e u^delta / delta!, nu ordered (000), (100), (010), (001), (200), (110), (101), ...
        q0, q1, q2 = mom[0], mom[1:4], mom[4:10]
        q += q0
        d += q1 + c * q0
        for (a, b), i in pos.items():
            w = 2.0 * q2[i] if a == b else q2[i]
            s[i] += w + c[a] * q1[b] + c[b] * q1[a] + c[a] * c[b] * q0
    assert pairs == plan.npairs
    assert abs(q - charge) <= 1e-12 * charge
    assert np.max(np.abs(d - dipole)) <= 1e-12 * max(1.0, np.max(np.abs(dipole)))
    # only the traceless part is fixed: a Gaussian's spread adds a multiple of
    # the identity to the second moments, which leaves the potential outside
    # unchanged (laplacian 1/r = 0)
    T, Tref = _traceless(s), _traceless(second)
    assert np.max(np.abs(T - Tref)) <= 1e-12 * np.max(np.abs(Tref))


@pytest.mark.parametrize("name", NAMES)
def test_multipole_limit_at_1000_bohr(name):
    m = molecule(name)
    electrons, dipole, s = moments.electron_moments(m.data)
    second = np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
    worst = quad = 0.0
    for direction in ([1, 0, 0], [0, -1, 0], [0, 0, 1], [0.6, 0.48, -0.64]):
        r = 1000.0 * np.asarray(direction, float) / np.linalg.norm(direction)
        d = np.linalg.norm(r)
        qt = (3.0 * r @ second @ r - d * d * np.trace(second)) / (2.0 * d ** 5)
        quad = max(quad, abs(qt))
        expansion = electrons / d + r @ dipole / d ** 3 + qt
        nuclear = sum(z / np.linalg.norm(r - a) for z, a in zip(m.data.charges, m.data.xyz))
        electronic = m.esp(r[None, :])[0] - nuclear
        worst = max(worst, abs(electronic + expansion))
    assert worst < 1e-2 * quad, (worst, quad)


@pytest.mark.parametrize("name", ["water", "o2", "acrolein"])
def test_evaluator_default_tolerance(name):
    m = molecule(name)
    grid = {"origin": np.array([-3.1, -2.9, -3.3]), "steps": np.array([[0.61, 0.05, 0], [0, 0.61, 0], [0, 0, 0.61]]),
            "counts": np.array([9, 8, 11])}
    exact = m.esp(oracle.grid_points(grid))
    a = m.evaluate(ESP, grid=grid)[0][0][0]
    b = m.evaluate(ESP, grid=grid, block_size=5)[0][0][0]
    assert np.array_equal(a, b)
    assert np.max(np.abs(a - exact)) <= 1e-8
    rng = np.random.default_rng(4)
    pts = m.data.xyz[rng.integers(0, len(m.data.xyz), 400)] + rng.normal(size=(400, 3)) * 2.5
    v = m.evaluate(ESP, points=pts)[0][0][0]
    assert np.max(np.abs(v - m.esp(pts))) <= 1e-8
    # a grid through a nucleus: the point on it gets the potential at the nucleus
    nuc = m.data.xyz[-1]
    g3 = {"origin": nuc - 0.5, "steps": np.eye(3) * 0.5, "counts": np.array([3, 3, 3])}
    v = m.evaluate(ESP, grid=g3)[0][0][0]
    assert abs(v[13] - m.esp(nuc[None, :])[0]) <= 1e-8


def test_pople20_shell_points_within_tolerance():
    """The 20-atom 6-31G*-shaped molecule with points on its 1.4 x vdW shell
    (the ESP benchmark's workload, sampled)."""
    m = molecule("pople20")
    pts = synthetic.vdw_shell_points(m.data.atomic_numbers, m.data.xyz, 3000)
    exact = m.esp(pts)
    for tol in (1e-6, 1e-8):
        v = m.evaluate(ESP, points=pts, potential_tolerance=tol)[0][0][0]
        assert np.max(np.abs(v - exact)) <= tol, tol


def test_far_field_rejects_invalid_use():
    m = molecule("water")
    pot = m.potential()
    with pytest.raises(ValueError, match="tolerance must be finite and nonnegative"):
        pot.plan(1.2, -1.0)
    with pytest.raises(ValueError, match="radius must be finite"):
        pot.plan(float("nan"), 1e-8)
    plan = pot.plan(1.2, 1e-8)
    c = m.data.xyz[0]
    with pytest.raises(ValueError, match="outside the planned radius"):
        pot.cluster(c[None, :] + [0.0, 0.0, 1.3], c, plan)
    other = molecule("o2").potential().plan(1.2, 1e-8)
    with pytest.raises(ValueError, match="does not match the potential"):
        pot.cluster(c[None, :], c, other)
    with pytest.raises(ValueError, match="potential tolerance must be finite and nonnegative"):
        m.evaluate(ESP, points=c[None, :], potential_tolerance=-1e-9)
