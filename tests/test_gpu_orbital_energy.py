"""Orbital-energy fields on the GPU (the weighted-sum mode of the AO-MO MFMA
kernel plus the field kernel) through the C-ABI: ALIE, LEA, LEAE, local
electronegativity and hardness, orbital-weighted Fukui functions f+, f-, f0
and the dual descriptor (src/orbital_energy_fields.hpp, src/evaluator.cpp:
481-520, 579-611, 676-682).

* Against the reference's external known answers: the Multiwfn cubes with the
  tolerances of tests/multiwfn_reference_test.cpp:395-430 and the truncation
  correction of its orbital-weighted comparison (:196-333).
* Against the oracle on the same grids and at scattered points, restricted
  (acrolein, water), unrestricted (O2) and synthetic molecules with more than
  128 MOs (several MO slot groups) and negative virtual energies: 1e-11
  relative to the largest magnitude of each field.
* The reference's errors (invalid_argument) for open shells, the spin
  components and a nonpositive Fukui delta.
"""
import numpy as np
import pytest

from kumo_hip.api import PropertyKind as K
from support import oracle
from support.fchk import FchkData
from support.reference_cases import (FCHK, ORBITAL_ENERGY, ORBITAL_WEIGHTED, check_multiwfn,
                                     check_orbital_weighted)

pytestmark = pytest.mark.gpu

ALL = [K.AverageLocalIonizationEnergy, K.LocalElectronAffinity, K.LocalElectronAttachmentEnergy,
       K.LocalElectronegativity, K.LocalHardness]
FUKUI = [K.OrbitalWeightedFukuiPlus, K.OrbitalWeightedFukuiMinus, K.OrbitalWeightedFukuiZero,
         K.OrbitalWeightedDualDescriptor]

_OR, _GPU = {}, {}


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


def close_to_oracle(got, ref, label, tol=1e-11):
    for g, r, p in zip(got, ref, range(len(got))):
        scale = np.abs(r).max() + 1e-300
        err = float(np.abs(g - r).max() / scale)
        assert err <= tol, f"{label} output {p}: {err:.3g} of the largest magnitude"


@pytest.mark.parametrize("case", ORBITAL_ENERGY, ids=[c[0] for c in ORBITAL_ENERGY])
def test_orbital_energy_fields_against_multiwfn_and_oracle(ctx, case):
    name = case[1]
    print(f"[gpu] orbital-energy field {case[0]}", flush=True)
    got = {}

    def ev(grid, props):
        g, _ = gmol(ctx, name).evaluate(props, grid=grid)
        got.update(g=g, grid=grid, props=props)
        return g
    check_multiwfn(case, ev, omol(name).data.xyz)
    ref, _ = omol(name).evaluate(got["props"], grid=got["grid"])
    close_to_oracle(got["g"], ref, case[0])


@pytest.mark.parametrize("case", ORBITAL_WEIGHTED, ids=[c[0] for c in ORBITAL_WEIGHTED])
def test_orbital_weighted_fukui_against_multiwfn_and_oracle(ctx, case):
    name = case[1]
    print(f"[gpu] orbital-weighted Fukui {case[0]}", flush=True)
    d = omol(name).data

    def ev(grid, props):
        return gmol(ctx, name).evaluate(props, grid=grid)[0]
    for st in check_orbital_weighted(case, ev, np.asarray(d.e_alpha), d.nalpha):
        ref, _ = omol(name).evaluate(st["props"], grid=st["grid"])
        close_to_oracle(st["outputs"], ref, st["label"])


@pytest.mark.parametrize("name", ["acrolein", "o2", "water"])
def test_points_match_oracle(ctx, name):
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(1000, 3)) * 2.5
    kinds = ALL + (FUKUI if name != "o2" else [])
    props = [(k, 0, 0, 0) for k in kinds] + [(K.Density, 0, 0, 0)]
    print(f"[gpu] orbital-energy fields at points, {name}", flush=True)
    g, _ = gmol(ctx, name).evaluate(props, points=pts)
    o, _ = omol(name).evaluate(props, points=pts)
    close_to_oracle(g, o, name)


def test_many_mos_and_negative_virtuals(ctx, tmp_path):
    """A 20-atom 6-31G*-shaped closed shell with about 200 MOs (two 128-slot
    groups) and random orbital energies in [-20, 2] (virtual MOs below zero
    enter LEAE), on a grid and at points, with block sizes 128 and 37."""
    from kumo_hip import synthetic
    from kumo_hip.binding import Molecule
    path, smol, _ = synthetic.cached_fchk(lambda: (synthetic.pople_molecule(), None), "pople20_seed3",
                                          directory=str(tmp_path))
    d = FchkData(path)
    assert d.nalpha == d.nbeta and len(d.e_alpha) > 128
    om = oracle.Molecule(d)
    gm = Molecule(ctx, fchk_path=path)
    try:
        c = np.mean(d.xyz, axis=0)
        grid = {"origin": c - 6.0, "steps": np.eye(3) * 0.6, "counts": [21, 20, 19]}
        props = [(k, 0, 0, 0) for k in ALL + FUKUI]
        for bs in (128, 37):
            print(f"[gpu] synthetic 6-31G* orbital-energy fields, block {bs}", flush=True)
            g, _ = gm.evaluate(props, grid=grid, block_size=bs)
            o, _ = om.evaluate(props, grid=grid, block_size=bs)
            close_to_oracle(g, o, f"synthetic block {bs}")
        leae = g[2][0]
        assert np.count_nonzero(leae) > 0.9 * leae.size
        pts = np.random.default_rng(2).normal(size=(500, 3)) * 3.0 + c
        g, _ = gm.evaluate(props, points=pts)
        o, _ = om.evaluate(props, points=pts)
        close_to_oracle(g, o, "synthetic points")
    finally:
        gm.close()


def test_unrestricted_synthetic(ctx, tmp_path):
    """Unrestricted: a water-like doublet (5 alpha, 4 beta), occupied and
    virtual MOs of both channels."""
    from kumo_hip import synthetic
    from kumo_hip.binding import Molecule

    def doublet():
        rng = np.random.default_rng(4)
        xyz = np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [-0.5, 1.7, 0.1]])
        recipe = lambda z: synthetic.POPLE_HEAVY if z > 1 else synthetic.POPLE_H  # noqa: E731
        return synthetic.SyntheticMolecule(rng, [8, 1, 1], xyz, recipe, 5, 4, True), None
    path, smol, _ = synthetic.cached_fchk(doublet, "water_doublet_seed4", directory=str(tmp_path))
    d = FchkData(path)
    om = oracle.Molecule(d)
    gm = Molecule(ctx, fchk_path=path)
    try:
        pts = np.random.default_rng(8).normal(size=(700, 3)) * 2.0
        props = [(k, 0, 0, 0) for k in ALL]
        print(f"[gpu] synthetic water-like orbital-energy fields (unrestricted={gm.info['unrestricted']})",
              flush=True)
        g, _ = gm.evaluate(props, points=pts)
        o, _ = om.evaluate(props, points=pts)
        close_to_oracle(g, o, "water-like")
    finally:
        gm.close()


def test_errors_as_the_reference(ctx):
    o2, acro = gmol(ctx, "o2"), gmol(ctx, "acrolein")
    p = np.zeros((1, 3))
    with pytest.raises(ValueError, match="the orbital-weighted Fukui function or dual descriptor is defined only "
                                         "for a restricted closed-shell calculation"):
        o2.evaluate([(K.OrbitalWeightedDualDescriptor, 0, 0, 0)], points=p)
    with pytest.raises(ValueError, match="fukui_delta must be finite and positive"):
        acro.evaluate([(K.AverageLocalIonizationEnergy, 0, 0, 0)], points=p, fukui_delta=-1.0)
    with pytest.raises(ValueError, match="the local hardness is defined only for the total density"):
        acro.evaluate([(K.LocalHardness, 2, 0, 0)], points=p)
    # delta changes f+ but not ALIE
    a1, _ = acro.evaluate([(K.OrbitalWeightedFukuiPlus, 0, 0, 0), (K.AverageLocalIonizationEnergy, 0, 0, 0)],
                          points=np.ones((5, 3)), fukui_delta=0.05)
    a2, _ = acro.evaluate([(K.OrbitalWeightedFukuiPlus, 0, 0, 0), (K.AverageLocalIonizationEnergy, 0, 0, 0)],
                          points=np.ones((5, 3)), fukui_delta=0.3)
    assert not np.allclose(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
    r, _ = omol("acrolein").evaluate([(K.OrbitalWeightedFukuiPlus, 0, 0, 0)], points=np.ones((5, 3)),
                                     fukui_delta=0.05)
    close_to_oracle([a1[0]], r, "delta 0.05")
