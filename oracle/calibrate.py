"""Calibration of the CPU oracle (test infrastructure) against the reference's
published CPU timings, so that the CPU baseline the benchmark reports is
shown not to be an unfairly slow stand-in.

Workloads (the reference's own benchmarks, HBC = hexa-peri-hexabenzocoronene):
  * docs/OPENMP.md:96-106 -- the Evaluator on HBC, Grid::around the atoms
    with a 3 bohr margin at 0.4 bohr (97 280 points): Hessian and AO_MO
    density; best of three runs, 1 and 8 threads (Apple M3, 4P+4E).
  * RELEASE_NOTES.md:825-836 -- HBC def2-TZVP, 50 752 points (the count of
    Grid::around with a 3 bohr margin at 0.5 bohr): AO_MO density and AO_MO
    Hessian + Laplacian + tau, 8 threads.
Both are evaluated as grids (Evaluator::evaluate(grid, request)).

Usage: python oracle/calibrate.py HBC_FCHK [HBC_FCHK ...] > summary.md
(the FCHK files are tests/data/hbc/*.fchk.gz of the reference tree). Timing
is of evaluate() only (no FCHK parsing), minimum of three runs after one
warm-up, as the reference times its benchmarks.
"""
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "chem-kumo_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

from kumo_hip.api import Backend, PropertyKind as K  # noqa: E402
from support import oracle  # noqa: E402
from support.fchk import FchkData  # noqa: E402

# Published seconds (docs/OPENMP.md:99-106; RELEASE_NOTES.md:831-836).
SPACING = {97280: 0.4, 50752: 0.5}
PUBLISHED = {
    ("def2svp", "hessian", 97280): {1: 2.688, 8: 0.632},
    ("def2svp", "ao_mo_density", 97280): {1: 0.416, 8: 0.088},
    ("def2tzvp", "hessian", 97280): {1: 5.779, 8: 1.386},
    ("def2tzvp", "ao_mo_density", 97280): {1: 0.837, 8: 0.188},
    ("def2tzvp", "ao_mo_density", 50752): {8: 0.09},
    ("def2tzvp", "ao_mo_hessian_lap_tau", 50752): {8: 0.73},
}

REQUESTS = {
    "hessian": ([(K.Hessian, 0, 0, 0)], Backend.AUTO),
    "ao_mo_density": ([(K.Density, 0, 0, 0)], Backend.AO_MO),
    "ao_mo_hessian_lap_tau": ([(K.Hessian, 0, 0, 0), (K.Laplacian, 0, 0, 0), (K.Tau, 0, 0, 0)], Backend.AO_MO),
}


def grid_around(xyz, margin, spacing):
    """Grid::around (src/grid.cpp:15-38)."""
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    counts = [max(2, int(np.ceil((hi[k] - lo[k] + 2 * margin) / spacing) + 1)) for k in range(3)]
    return {"origin": lo - margin, "steps": np.eye(3) * spacing, "counts": np.array(counts, np.int32)}


def best_of(fn, runs=3):
    fn()
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return min(times), float(np.median(times))


def cpu_description():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    lscpu = ""
    try:
        lscpu = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        pass
    fields = {}
    for line in lscpu.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    return model, fields


def main(paths):
    model, fields = cpu_description()
    lib = oracle.lib()
    threads_max = int(lib.orc_threads())
    print("# CPU oracle calibration against the reference's published HBC timings\n")
    print(f"CPU: {model}; lscpu: sockets {fields.get('Socket(s)', '?')}, cores per socket "
          f"{fields.get('Core(s) per socket', '?')}, threads per core {fields.get('Thread(s) per core', '?')}, "
          f"online CPUs {fields.get('On-line CPU(s) list', '?')}. Oracle threads used: 1 and {threads_max}.\n")
    print("Reference figures: Apple M3 (4 performance + 4 efficiency cores), MacPorts Clang 22, -O3, OpenMP "
          "(docs/OPENMP.md:96-106, RELEASE_NOTES.md:825-836).\n")
    print("| Basis | Request | Points | Threads | Oracle s (min / median of 3) | Published s | Oracle / published |")
    print("|---|---|---:|---:|---:|---:|---:|")
    for path in paths:
        basis = "def2tzvp" if "tzvp" in os.path.basename(path) else "def2svp"
        data = FchkData(path)
        mol = oracle.Molecule(data)
        for (b, req, npts), pub in PUBLISHED.items():
            if b != basis:
                continue
            grid = grid_around(np.asarray(data.xyz), 3.0, SPACING[npts])
            count = int(np.prod(grid["counts"]))
            props, backend = REQUESTS[req]
            for th in sorted(pub):
                use = th if th <= threads_max else threads_max
                lib.orc_set_threads(use)
                tmin, tmed = best_of(lambda: mol.evaluate(props, grid=grid, backend=backend))
                print(f"| {basis} | {req} | {count} | {use} | {tmin:.3f} / {tmed:.3f} | {pub[th]:.3f} ({th} thr) | "
                      f"{tmin / pub[th]:.2f} |", flush=True)
        lib.orc_set_threads(threads_max)
        print(f"\n{basis}: {data.nbasis} AOs; grids Grid::around(3.0, 0.4) and (3.0, 0.5).\n")


if __name__ == "__main__":
    main(sys.argv[1:])
