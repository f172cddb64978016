"""Gaussian Cube reader for the fixtures (plain or gzip)."""
import gzip

import numpy as np


def read_cube(path):
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as fh:
        text = fh.read().split("\n")
    parts = text[2].split()
    natoms = int(parts[0])
    origin = np.array([float(v) for v in parts[1:4]])
    counts = []
    steps = np.zeros((3, 3))  # column a = step vector of axis a
    for a in range(3):
        p = text[3 + a].split()
        counts.append(int(p[0]))
        steps[:, a] = [float(v) for v in p[1:4]]
    atoms = []
    for a in range(abs(natoms)):
        p = text[6 + a].split()
        atoms.append((int(p[0]), float(p[1]), float(p[2]), float(p[3]), float(p[4])))
    start = 6 + abs(natoms)
    orbital = None
    if natoms < 0:
        p = text[start].split()
        orbital = int(p[1])
        start += 1
    values = np.array(" ".join(text[start:]).split(), dtype=np.float64)
    assert values.size == counts[0] * counts[1] * counts[2], "truncated cube"
    return {"origin": origin, "steps": steps, "counts": np.array(counts, dtype=np.int32),
            "atoms": atoms, "values": values, "orbital": orbital}


def recover_standard_step(steps):
    """Gaussian's -3 grid prints 1/6 bohr as 0.166667 (tests/cube_reference_test.cpp:24-38)."""
    s = steps.copy()
    h = 1.0 / 6.0
    for a in range(3):
        if abs(s[a, a] - h) < 0.5e-6 + 1e-12:
            s[a, a] = h
    return s
