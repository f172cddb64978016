"""The reference's external known-answer tests, restated over the committed
fixtures so that the same checks run against the oracle (CPU) and the HIP
path (GPU). Tolerances and scales follow the reference tests:
tests/cube_reference_test.cpp:146-160, tests/orbital_reference_test.cpp,
tests/multiwfn_reference_test.cpp:37-236, 335-487."""
import math
import os

import numpy as np

from kumo_hip.api import DensityComponent as DC, PropertyKind as K, channels
from support.cube import read_cube, recover_standard_step

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "golden")
FCHK = {"o2": os.path.join(GOLDEN, "fchk", "o2.fchk.gz"),
        "acrolein": os.path.join(GOLDEN, "fchk", "acrolein.fchk.gz"),
        "water": os.path.join(GOLDEN, "fchk", "h2o_b3lyp_def2tzvp.fchk.gz")}

REL = 2e-5
# name -> (molecule, cube, kind, component, tolerance{relative, absolute, exclusion, scale}, axis)
MULTIWFN = [
    ("water ELF", "water", "h2o_elf", K.Elf, DC.Total, {}, 0),
    ("water RDG", "water", "h2o_rdg", K.ReducedGradient, DC.Total, {}, 0),
    ("water sign(lambda2) rho", "water", "h2o_sign_lambda2_rho", K.SignLambda2Density, DC.Total, {}, 0),
    ("O2 ELF", "o2", "o2_elf", K.Elf, DC.Total, {}, 0),
    ("water ESP", "water", "h2o_esp", K.ElectrostaticPotential, DC.Total, {"absolute": 1e-7, "exclusion": 1e-3}, 0),
    ("O2 ESP", "o2", "o2_esp", K.ElectrostaticPotential, DC.Total, {"absolute": 1e-7, "exclusion": 1e-3}, 0),
    ("water LOL", "water", "h2o_lol", K.Lol, DC.Total, {}, 0),
    ("O2 LOL", "o2", "o2_lol", K.Lol, DC.Total, {}, 0),
    ("water IRI", "water", "h2o_iri", K.InteractionRegionIndicator, DC.Total, {"relative": 5e-5}, 0),
    ("water DORI", "water", "h2o_dori", K.DensityOverlapRegionsIndicator, DC.Total, {"absolute": 1e-5}, 0),
    ("water K", "water", "h2o_kinetic_k", K.HamiltonianKineticEnergy, DC.Total, {"relative": 5e-5, "scale": "energy"}, 0),
    ("O2 K", "o2", "o2_kinetic_k", K.HamiltonianKineticEnergy, DC.Total, {"relative": 5e-5, "scale": "energy"}, 0),
    ("water V", "water", "h2o_potential_energy", K.PotentialEnergyDensity, DC.Total, {"relative": 5e-5, "scale": "energy"}, 0),
    ("water H", "water", "h2o_energy_density", K.EnergyDensity, DC.Total, {"relative": 5e-5, "scale": "energy"}, 0),
    ("O2 spin polarization", "o2", "o2_spin_polarization", K.SpinPolarization, DC.Difference, {"absolute": 1e-6}, 0),
    ("water Shannon", "water", "h2o_shannon", K.ShannonEntropyDensity, DC.Total, {"scale": "shannon"}, 0),
    ("water Fisher", "water", "h2o_fisher", K.FisherInformationDensity, DC.Total, {"relative": 1e-4}, 0),
    ("water ELF-TS", "water", "h2o_elf_ts", K.ElfTsirelsonStash, DC.Total, {}, 0),
    ("water LOL-TS", "water", "h2o_lol_ts", K.LolTsirelsonStash, DC.Total, {}, 0),
    ("water G", "water", "h2o_kinetic_g", K.Tau, DC.Total, {}, 0),
    ("water G_x", "water", "h2o_kinetic_g_x", K.TauComponents, DC.Total, {"relative": 5e-5}, 0),
    ("water K_x", "water", "h2o_kinetic_k_x", K.HamiltonianKineticEnergyComponents, DC.Total,
     {"relative": 5e-5, "scale": "energy_axis"}, 0),
    ("water |V|/G", "water", "h2o_potential_kinetic_ratio", K.PotentialKineticRatio, DC.Total,
     {"relative": 5e-5, "scale": "virial"}, 0),
    ("water H/rho", "water", "h2o_energy_density_per_electron", K.EnergyDensityPerElectron, DC.Total,
     {"relative": 5e-5, "scale": "per_electron"}, 0),
    ("water G/rho", "water", "h2o_kinetic_g_per_electron", K.KineticEnergyPerElectron, DC.Total, {}, 0),
]


# Orbital-energy fields (tests/multiwfn_reference_test.cpp:395-430): ALIE, LEA
# and the local hardness take the default tolerance; LEAE adds 1e-10 hartree
# absolute and must hold the relative tolerance alone on at least 10000
# points with |LEAE| >= 1e-3; the local electronegativity is compared on the
# scale (|ALIE| + |LEA|) / 2.
_LEAE = {"absolute": 1e-10, "significant": 1e-3, "minimum_significant": 10000}
ORBITAL_ENERGY = [
    ("acrolein ALIE", "acrolein", "acrolein_alie", K.AverageLocalIonizationEnergy, DC.Total, {}, 0),
    ("O2 ALIE", "o2", "o2_alie", K.AverageLocalIonizationEnergy, DC.Total, {}, 0),
    ("acrolein LEA", "acrolein", "acrolein_lea", K.LocalElectronAffinity, DC.Total, {}, 0),
    ("O2 LEA", "o2", "o2_lea", K.LocalElectronAffinity, DC.Total, {}, 0),
    ("acrolein LEAE", "acrolein", "acrolein_leae", K.LocalElectronAttachmentEnergy, DC.Total, _LEAE, 0),
    ("O2 LEAE", "o2", "o2_leae", K.LocalElectronAttachmentEnergy, DC.Total, _LEAE, 0),
    ("acrolein local electronegativity", "acrolein", "acrolein_local_electronegativity", K.LocalElectronegativity,
     DC.Total, {"scale": "orbital"}, 0),
    ("acrolein local hardness", "acrolein", "acrolein_local_hardness", K.LocalHardness, DC.Total, {}, 0),
]
# Orbital-weighted Fukui functions (tests/multiwfn_reference_test.cpp:196-333,
# 485-486): (label, molecule, Cube prefix, also f0 and the dual descriptor).
ORBITAL_WEIGHTED = [("acrolein", "acrolein", "acrolein", True), ("water", "water", "h2o", False)]


def cube_grid(cube):
    return {"origin": cube["origin"], "steps": cube["steps"], "counts": cube["counts"]}


def grid_xyz(grid):
    counts = grid["counts"]
    i, j, k = np.meshgrid(np.arange(counts[0]), np.arange(counts[1]), np.arange(counts[2]), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(float)
    return grid["origin"] + ijk @ np.asarray(grid["steps"]).T


def check_multiwfn(case, evaluate, atoms_xyz):
    """evaluate(grid, props) -> list of arrays (channels x n). Returns a dict of
    statistics and raises AssertionError like the reference test."""
    label, mol, cube_name, kind, comp, tol, axis = case
    cube = read_cube(os.path.join(GOLDEN, "multiwfn", cube_name + ".cube.gz"))
    grid = cube_grid(cube)
    props = [(int(kind), int(comp), 0, 0), (K.Density, 0, 0, 0), (K.Tau, 0, 0, 0), (K.Laplacian, 0, 0, 0),
             (K.TauComponents, 0, 0, 0), (K.Hessian, 0, 0, 0)]
    if tol.get("scale") == "orbital":
        props += [(K.AverageLocalIonizationEnergy, 0, 0, 0), (K.LocalElectronAffinity, 0, 0, 0)]
    outs = evaluate(grid, props)
    actual = outs[0][axis]
    rho, tau, lap = outs[1][0], outs[2][0], outs[3][0]
    tau_axis = outs[4][axis]
    hess_axis = outs[5][[0, 3, 5][axis]]
    expected = cube["values"]
    rel = tol.get("relative", REL)
    ab = tol.get("absolute", 0.0)
    excl = tol.get("exclusion", 0.0)
    scale_kind = tol.get("scale", "value")
    xyz = grid_xyz(grid)
    keep = np.ones(expected.size, bool)
    if excl > 0:
        for a in atoms_xyz:
            keep &= np.linalg.norm(xyz - a, axis=1) >= excl
    with np.errstate(divide="ignore", invalid="ignore"):
        if scale_kind == "energy":
            scale = np.abs(tau) + 0.25 * np.abs(lap)
        elif scale_kind == "shannon":
            scale = np.where(rho > 0, rho * (1 + np.abs(np.log(np.where(rho > 0, rho, 1.0)))), 0.0)
        elif scale_kind == "energy_axis":
            scale = np.abs(tau_axis) + 0.25 * np.abs(hess_axis)
        elif scale_kind == "per_electron":
            scale = np.where(rho > 0, (np.abs(tau) + 0.25 * np.abs(lap)) / np.where(rho > 0, rho, 1.0), 0.0)
        elif scale_kind == "virial":
            scale = np.where(tau > 0, (2 * np.abs(tau) + 0.25 * np.abs(lap)) / np.where(tau > 0, tau, 1.0), 0.0)
        elif scale_kind == "orbital":
            scale = 0.5 * (np.abs(outs[6][0]) + np.abs(outs[7][0]))
        else:
            scale = np.abs(expected)
    diff = np.abs(actual - expected)
    allowed = rel * scale + ab
    ratio = (diff / np.maximum(allowed, 1e-300))[keep]
    worst = float(ratio.max())
    sig = np.abs(expected) > ab
    same_sign = bool(np.all(((actual < 0) == (expected < 0))[keep & sig]))
    assert same_sign, f"{label}: sign differs from the reference"
    assert worst <= 1.0, f"{label}: max difference {worst:.3g} of the tolerance"
    if tol.get("significant"):
        big = keep & (np.abs(expected) >= tol["significant"])
        assert big.sum() >= tol["minimum_significant"], f"{label}: {big.sum()} significant points"
        rel_big = float((diff[big] / np.abs(expected[big])).max())
        assert rel_big <= rel, f"{label}: {rel_big:.3g} relative where |value| >= {tol['significant']}"
    return {"label": label, "worst": worst, "points": int(keep.sum()), "outputs": outs}


def check_orbital_weighted(case, evaluate, energies, homo):
    """compare_orbital_weighted (tests/multiwfn_reference_test.cpp:196-333):
    Multiwfn keeps only the orbitals whose unnormalized Gaussian
    exp(-((mu - e_i) / 0.1)^2) is at least 1e-3, scanning up from the LUMO (f+)
    and down from the HOMO (f-), and normalizes over those; the truncated sums
    are rebuilt from the evaluated orbitals and their difference to the
    evaluator's f+ / f- is added to the Cube values. Then: 2e-5 relative +
    1e-8 absolute everywhere, the relative tolerance alone where |value| >=
    1e-4 (at least 5000 points). evaluate(grid, props) -> list of arrays.
    homo: occupied alpha orbitals; energies: alpha orbital energies."""
    label, mol, prefix, everything = case
    delta, cut, absolute, significant_value, minimum = 0.1, 1e-3, 1e-8, 1e-4, 5000
    count = len(energies)
    mu = 0.5 * (energies[homo - 1] + energies[homo])

    def gaussian(i):  # one-based
        x = (mu - energies[i - 1]) / delta
        return math.exp(-x * x)
    plus_kept, minus_kept = [], []
    i = homo + 1
    while i <= count and gaussian(i) >= cut:
        plus_kept.append(i)
        i += 1
    i = homo
    while i >= 1 and gaussian(i) >= cut:
        minus_kept.append(i)
        i -= 1
    assert plus_kept and minus_kept, f"{label}: Multiwfn keeps at least the LUMO and the HOMO"
    functions = [("ow_fukui_plus", K.OrbitalWeightedFukuiPlus), ("ow_fukui_minus", K.OrbitalWeightedFukuiMinus)]
    if everything:
        functions += [("ow_fukui_zero", K.OrbitalWeightedFukuiZero),
                      ("ow_dual_descriptor", K.OrbitalWeightedDualDescriptor)]
    stats = []
    for name, kind in functions:
        cube = read_cube(os.path.join(GOLDEN, "multiwfn", f"{prefix}_{name}.cube.gz"))
        grid = cube_grid(cube)
        expected = cube["values"]
        props = [(int(kind), 0, 0, 0), (K.OrbitalWeightedFukuiPlus, 0, 0, 0), (K.OrbitalWeightedFukuiMinus, 0, 0, 0)]
        props += [(K.Orbital, 0, 0, i) for i in plus_kept + minus_kept]
        outs = evaluate(grid, props)
        actual, plus, minus = outs[0][0], outs[1][0], outs[2][0]
        amp = {i: outs[3 + q][0] for q, i in enumerate(plus_kept + minus_kept)}

        def truncated(kept):
            s = sum(gaussian(i) * amp[i] ** 2 for i in kept)
            return s / sum(gaussian(i) for i in kept)
        d_plus = plus - truncated(plus_kept)
        d_minus = minus - truncated(minus_kept)
        correction = {K.OrbitalWeightedFukuiPlus: d_plus, K.OrbitalWeightedFukuiMinus: d_minus,
                      K.OrbitalWeightedFukuiZero: 0.5 * (d_plus + d_minus),
                      K.OrbitalWeightedDualDescriptor: d_plus - d_minus}[kind]
        scale = np.abs(plus) + np.abs(minus) if kind == K.OrbitalWeightedDualDescriptor else np.abs(actual)
        diff = np.abs(actual - (expected + correction))
        worst = float((diff / (REL * scale + absolute)).max())
        big = np.abs(expected) >= significant_value
        worst_big = float((diff[big] / np.maximum(REL * scale[big], 1e-300)).max())
        assert big.sum() >= minimum, f"{label} {name}: {big.sum()} points with |value| >= {significant_value}"
        assert worst_big <= 1.0, f"{label} {name}: {worst_big:.3g} of the relative tolerance where significant"
        assert worst <= 1.0, f"{label} {name}: {worst:.3g} of the tolerance"
        stats.append({"label": f"{label} {name}", "worst": worst, "worst_significant": worst_big, "outputs": outs,
                      "grid": grid, "props": props})
    return stats


DENSITY_CUBES = ["o2_density", "o2_alpha", "o2_beta", "o2_spin", "acrolein_density"]
MO_CUBES = ["o2_alpha_mo_3", "o2_alpha_mo_9", "o2_alpha_mo_10", "o2_beta_mo_3", "o2_beta_mo_7", "o2_beta_mo_8",
            "acrolein_mo_13", "acrolein_mo_15", "acrolein_mo_16", "acrolein_mo_18"]


def load_sample(name):
    z = np.load(os.path.join(GOLDEN, "cubegen", name + ".npz"))
    d = {k: z[k] for k in z.files}
    d["steps"] = recover_standard_step(d["steps"])
    d["grid"] = {"origin": d["origin"], "steps": d["steps"], "counts": d["counts"]}
    return d


def check_density_cube(name, values_full):
    """values_full: the density on the whole Cube grid (Field order)."""
    d = load_sample(name)
    idx = d["index"]
    ref = d["value"]
    got = values_full[idx]
    diff = np.abs(got - ref)
    max_abs = float(diff.max())
    rel_l1 = float(diff.sum() / np.abs(ref).sum())
    volume = abs(np.linalg.det(d["steps"]))
    count_diff = abs(float(d["total"]) * volume - float(values_full.sum()) * volume)
    assert max_abs <= 2e-3, f"{name}: max_abs {max_abs}"
    assert rel_l1 <= 1e-5, f"{name}: relative L1 {rel_l1}"
    assert count_diff <= 0.5, f"{name}: electron count differs by {count_diff}"
    return {"max_abs": max_abs, "rel_l1": rel_l1, "electrons": float(values_full.sum()) * volume}


def check_mo_cube(name, values_full):
    d = load_sample(name)
    idx = d["index"]
    ref = d["value"]
    got = values_full[idx]
    if np.dot(got, ref) < 0:
        got = -got
    diff = np.abs(got - ref)
    max_abs = float(diff.max())
    rel_l1 = float(diff.sum() / np.abs(ref).sum())
    assert max_abs <= 2e-4, f"{name}: max_abs {max_abs}"
    assert rel_l1 <= 2e-5, f"{name}: relative L1 {rel_l1}"
    return {"max_abs": max_abs, "rel_l1": rel_l1}
