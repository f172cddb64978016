"""Cuts the head of the Gaussian cubegen density Cubes of the reference's test
data (tests/data/o2_density.cube, tests/data/acrolein_density.cube; provenance
in tests/REFERENCE_CUBES.md there): the header and the sample lines of the
first three runs along the third axis (cubegen starts six samples per line
afresh in every run). The excerpts pin the text of one sample (width 13, six
significant digits, two-digit exponents) and the line structure of the first
run, which the Cube writer shares with cubegen.

Usage: python3 make_excerpts.py [/root/reference/tests/data]
"""
import math
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data"
HERE = os.path.dirname(os.path.abspath(__file__))

for name in ("o2_density.cube", "acrolein_density.cube"):
    with open(os.path.join(SRC, name)) as fh:
        lines = fh.read().split("\n")
    natoms = abs(int(lines[2].split()[0]))
    nz = int(lines[5].split()[0])
    keep = 6 + natoms + 3 * math.ceil(nz / 6)
    with open(os.path.join(HERE, name.replace(".cube", "_head.txt")), "w") as out:
        out.write("\n".join(lines[:keep]) + "\n")
