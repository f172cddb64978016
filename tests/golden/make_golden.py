"""Builds the golden fixtures under tests/golden/ from the reference's own
test data (tests/data of chem-kumo 2.1.0). Run in a container where the
reference is mounted:  python tests/golden/make_golden.py [/root/reference]

* fchk/: the FCHK inputs, copied byte for byte (o2, acrolein, water def2-TZVP).
* multiwfn/: Multiwfn 2026.9.1 Cube files of derived fields and its surface
  analysis reports, copied byte for byte (provenance: the reference's
  tests/REFERENCE_CUBES.md).
* cubegen/*.npz: Gaussian 09 cubegen density and MO cubes, which are 3-10 MB of
  text each, reduced to the grid header, a deterministic sample of points
  (every STRIDE-th point plus every point whose magnitude exceeds KEEP) with
  their Cube values, and the sum of all values (for the electron count).
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from support.cube import read_cube  # noqa: E402

STRIDE = 29
KEEP = 0.5

FCHK = ["o2.fchk.gz", "acrolein.fchk.gz", "hybrids/h2o_b3lyp_def2tzvp.fchk.gz"]
MULTIWFN = ["h2o_elf", "h2o_rdg", "h2o_sign_lambda2_rho", "o2_elf", "h2o_esp", "o2_esp", "h2o_lol", "o2_lol",
            "h2o_iri", "h2o_dori", "h2o_kinetic_k", "o2_kinetic_k", "h2o_potential_energy", "h2o_energy_density",
            "o2_spin_polarization", "h2o_shannon", "h2o_fisher", "h2o_elf_ts", "h2o_lol_ts", "h2o_kinetic_g",
            "h2o_kinetic_g_x", "h2o_kinetic_k_x", "h2o_potential_kinetic_ratio", "h2o_energy_density_per_electron",
            "h2o_kinetic_g_per_electron",
            # orbital-energy fields (tests/multiwfn_reference_test.cpp:395-430, 485-486)
            "acrolein_alie", "o2_alie", "acrolein_lea", "o2_lea", "acrolein_leae", "o2_leae",
            "acrolein_local_electronegativity", "acrolein_local_hardness", "acrolein_ow_fukui_plus",
            "acrolein_ow_fukui_minus", "acrolein_ow_fukui_zero", "acrolein_ow_dual_descriptor", "h2o_ow_fukui_plus",
            "h2o_ow_fukui_minus"]
# Multiwfn quantitative surface analyses (tests/multiwfn_surface_reference_test.cpp, meson.build:746-782).
SURFACES = ["h2o_esp", "o2_esp", "acrolein_esp", "h2o_alie", "o2_alie", "acrolein_alie", "acrolein_leae", "o2_leae",
            "acrolein_lea"]
DENSITY = [("o2_density", "o2", 0), ("o2_alpha", "o2", 1), ("o2_beta", "o2", 2), ("o2_spin", "o2", 3),
           ("acrolein_density", "acrolein", 0)]
MO = [("o2_alpha_mo_3", "o2", 0, 3), ("o2_alpha_mo_9", "o2", 0, 9), ("o2_alpha_mo_10", "o2", 0, 10),
      ("o2_beta_mo_3", "o2", 1, 3), ("o2_beta_mo_7", "o2", 1, 7), ("o2_beta_mo_8", "o2", 1, 8),
      ("acrolein_mo_13", "acrolein", 0, 13), ("acrolein_mo_15", "acrolein", 0, 15),
      ("acrolein_mo_16", "acrolein", 0, 16), ("acrolein_mo_18", "acrolein", 0, 18)]


def sample(cube, out, **meta):
    v = cube["values"]
    idx = np.arange(0, v.size, STRIDE)
    big = np.nonzero(np.abs(v) > KEEP)[0]
    idx = np.union1d(idx, big).astype(np.int64)
    np.savez_compressed(out, origin=cube["origin"], steps=cube["steps"], counts=cube["counts"],
                        index=idx, value=v[idx], total=np.array(v.sum()), abs_total=np.array(np.abs(v).sum()),
                        npoints=np.array(v.size), **{k: np.array(x) for k, x in meta.items()})


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = os.path.join(ref, "tests", "data")
    for f in FCHK:
        shutil.copyfile(os.path.join(data, f), os.path.join(HERE, "fchk", os.path.basename(f)))
    for f in MULTIWFN:
        shutil.copyfile(os.path.join(data, "multiwfn", f + ".cube.gz"), os.path.join(HERE, "multiwfn", f + ".cube.gz"))
    for f in SURFACES:
        shutil.copyfile(os.path.join(data, "multiwfn", f + "_surface.txt"),
                        os.path.join(HERE, "multiwfn", f + "_surface.txt"))
    for name, mol, comp in DENSITY:
        sample(read_cube(os.path.join(data, name + ".cube")), os.path.join(HERE, "cubegen", name + ".npz"),
               molecule=mol, component=comp)
    for name, mol, channel, index in MO:
        sample(read_cube(os.path.join(data, "mo_reference_cubes", name + ".cube")),
               os.path.join(HERE, "cubegen", name + ".npz"), molecule=mol, channel=channel, orbital=index)


if __name__ == "__main__":
    main()
