"""Minimal Gaussian formatted-checkpoint reader for the tests, written
independently of the product's C++ reader so that the two cross-check.
Record layout follows the reference reader (src/fchk.cpp:120-190): a label
in the first 43 columns, the type letter (I or R), then "N=" and a count for
arrays. Only I and R records are kept."""
import gzip

import numpy as np


def _open(path):
    with open(path, "rb") as fh:
        head = fh.read(2)
    if head == b"\x1f\x8b":
        return gzip.open(path, "rt")
    return open(path, "rt")


def read_fchk(path):
    records = {}
    with _open(path) as fh:
        lines = fh.read().splitlines()
    i = 2
    while i < len(lines):
        line = lines[i]
        i += 1
        if len(line) < 44:
            continue
        label = line[:43].rstrip()
        meta = line[43:].split()
        if not label or not meta or meta[0] not in ("I", "R"):
            continue
        kind = meta[0]
        if len(meta) >= 3 and meta[1] == "N=":
            count = int(meta[2])
            tokens = []
            while len(tokens) < count:
                tokens.extend(lines[i].split())
                i += 1
            if kind == "I":
                records[label] = np.array([int(t) for t in tokens], dtype=np.int64)
            else:
                records[label] = np.array(
                    [float(t.replace("D", "E").replace("d", "E")) for t in tokens], dtype=np.float64)
        else:
            records[label] = int(meta[1]) if kind == "I" else float(meta[1].replace("D", "E"))
    return records


def unpack_lower(packed, n):
    """Row-wise packed lower triangle -> symmetric matrix (src/molecule.cpp:20-33)."""
    m = np.zeros((n, n))
    il = np.tril_indices(n)
    m[il] = packed
    m = m + m.T - np.diag(np.diag(m))
    return m


class FchkData:
    """The arrays the evaluator needs, in FCHK order."""

    def __init__(self, path):
        r = read_fchk(path)
        self.records = r
        self.atomic_numbers = r["Atomic numbers"].astype(np.int32)
        self.natoms = len(self.atomic_numbers)
        self.charges = np.asarray(r.get("Nuclear charges", self.atomic_numbers.astype(float)), dtype=float)
        self.xyz = r["Current cartesian coordinates"].reshape(-1, 3)
        self.shell_types = r["Shell types"].astype(np.int32)
        self.nprim = r["Number of primitives per shell"].astype(np.int32)
        self.shell_atom = r["Shell to atom map"].astype(np.int32)
        self.exponents = r["Primitive exponents"]
        self.coefficients = r["Contraction coefficients"]
        self.p_coefficients = r.get("P(S=P) Contraction coefficients")
        self.shell_xyz = r["Coordinates of each shell"].reshape(-1, 3)
        self.nbasis = int(r["Number of basis functions"])
        self.p_total = unpack_lower(r["Total SCF Density"], self.nbasis)
        if "Spin SCF Density" in r:
            self.p_spin = unpack_lower(r["Spin SCF Density"], self.nbasis)
        else:
            self.p_spin = np.zeros_like(self.p_total)
        self.nalpha = int(r["Number of alpha electrons"])
        self.nbeta = int(r["Number of beta electrons"])
        self.c_alpha = r.get("Alpha MO coefficients")
        self.c_beta = r.get("Beta MO coefficients")
        self.e_alpha = r.get("Alpha Orbital Energies")
        self.e_beta = r.get("Beta Orbital Energies")

    def mo_matrix(self, channel):
        """nbasis x nmo, one MO per column (src/mo_coefficients.hpp:21-34)."""
        c = self.c_beta if (channel == 1 and self.c_beta is not None) else self.c_alpha
        return c.reshape(-1, self.nbasis).T
