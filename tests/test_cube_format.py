"""Cube output without a GPU: the decimal formatter of the GPU Cube writer
compiled for the host against glibc's printf, and the oracle's restatement
of Field::write_cube (src/cube.cpp:70-111, 253-262) pinned by

 - the Gaussian cubegen fixtures of the reference's test data (excerpts in
   tests/golden/cube_text, cut by make_excerpts.py there): the header lines of
   the grid and the atoms, the text of every sample (width 13, six significant
   digits, two-digit exponents) and the lines of the first run along the
   third axis, which the reference's writer shares with cubegen, and the
   numbers of the header lines (six decimals; the field widths differ);
 - tests/atom_charge_cube_test.cpp: Z = 26 with charge 16, atom count 1 or
   -1, two points per axis;
 - tests/cube_io_test.cpp:57-69: a nonfinite sample is rejected (and the file
   is not opened), and src/cube.cpp:258-259 (orbital index 0 or above 9999);
 - the mathematics of the format: the samples read back within half a unit
   of the sixth significant digit.
The break after every sixth sample of the whole sequence (not of each run,
as cubegen) in the runs after the first follows src/cube.cpp:107 and has no
external fixture: the token sequence of those runs is still pinned.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from support import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXCERPTS = os.path.join(ROOT, "tests", "golden", "cube_text")


def test_formatter_host_build_matches_printf(tmp_path):
    exe = str(tmp_path / "sci_format_test")
    subprocess.check_call(["g++", "-std=c++20", "-O2", "-ffp-contract=off",
                           os.path.join(ROOT, "tests", "cpp", "sci_format_test.cpp"), "-o", exe])
    out = subprocess.run([exe, "200000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:]
    assert " 0 mismatches" in out.stdout


def parse_header(lines):
    natoms = int(lines[2].split()[0])
    origin = [float(v) for v in lines[2].split()[1:4]]
    counts, steps = [], np.zeros((3, 3))
    for a in range(3):
        p = lines[3 + a].split()
        counts.append(int(p[0]))
        steps[:, a] = [float(v) for v in p[1:4]]
    z, q, xyz = [], [], []
    for i in range(abs(natoms)):
        p = lines[6 + i].split()
        z.append(int(p[0]))
        q.append(float(p[1]))
        xyz.append([float(v) for v in p[2:5]])
    grid = {"origin": origin, "steps": steps, "counts": counts}
    return grid, (np.array(z), np.array(q), np.array(xyz)), abs(natoms)


@pytest.mark.parametrize("name", ["o2_density", "acrolein_density"])
def test_oracle_cube_matches_cubegen_excerpt(tmp_path, name):
    lines = open(os.path.join(EXCERPTS, name + "_head.txt")).read().split("\n")
    grid, atoms, natoms = parse_header(lines)
    nz = grid["counts"][2]
    first = 6 + natoms
    samples = np.array(" ".join(lines[first:]).split(), dtype=np.float64)
    assert samples.size == 3 * nz
    n = int(np.prod(grid["counts"]))
    values = np.zeros(n)
    values[:samples.size] = samples
    path = str(tmp_path / "o.cube")
    rc, msg = oracle.write_cube(path, grid, atoms, values)
    assert rc == 0, msg
    with open(path) as fh:
        out = [fh.readline().rstrip("\n") for _ in range(first + 3 * math.ceil(nz / 6) + 2)]
    assert out[0] == "chem-kumo 2.1.0" and out[1] == "Electron density"
    # Header: the same numbers with six decimals (cubegen's fields are 12
    # wide, the reference's 13, src/cube.cpp:85-98); cubegen appends the
    # number of values per point to the atom-count line.
    assert out[2].split() == lines[2].split()[:4] and lines[2].split()[4:] == ["1"]
    assert [ln.split() for ln in out[3:first]] == [ln.split() for ln in lines[3:first]]
    assert all(len(ln) == 5 + 13 * (3 if i < 6 else 4) for i, ln in enumerate(out[2:first], start=2))
    run = math.ceil(nz / 6)
    assert out[first:first + run] == [s.lower() for s in lines[first:first + run]]
    # Later runs: the same tokens, as the text of each sample.
    got = " ".join(out[first:first + 3 * run]).split()[:3 * nz]
    want = [t.lower() for t in " ".join(lines[first:]).split()]
    assert got == want


def test_oracle_atom_charge_case(tmp_path):
    # tests/atom_charge_cube_test.cpp: an iron atom (Z 26) with ECP charge 16.
    grid = {"origin": [0, 0, 0], "steps": np.eye(3), "counts": [2, 2, 2]}
    atoms = (np.array([26]), np.array([16.0]), np.zeros((1, 3)))
    for orbital in (None, 1):
        path = str(tmp_path / "c.cube")
        rc, msg = oracle.write_cube(path, grid, atoms, np.full(8, 0.5), orbital_index=orbital)
        assert rc == 0, msg
        lines = open(path).read().split("\n")
        head = lines[2].split()
        assert int(head[0]) == (-1 if orbital else 1)
        for a in range(3):
            assert int(lines[3 + a].split()[0]) == 2
        z, q = lines[6].split()[:2]
        assert int(z) == 26 and abs(float(q) - 16.0) < 1e-6
        body = lines[7:]
        if orbital:
            assert lines[7] == "    1    1"
            body = lines[8:]
        # Two samples per run of the third axis, one run per line.
        assert body == ["  5.00000e-01  5.00000e-01"] * 4 + [""]


def test_oracle_rejects_before_opening(tmp_path):
    grid = {"origin": [0, 0, 0], "steps": np.eye(3), "counts": [2, 2, 2]}
    atoms = (np.array([1]), np.array([1.0]), np.zeros((1, 3)))
    for bad in (np.nan, np.inf, -np.inf):
        v = np.ones(8)
        v[5] = bad
        path = tmp_path / "bad.cube"
        rc, msg = oracle.write_cube(str(path), grid, atoms, v)
        assert rc == 1 and msg == "nonfinite Cube sample"
        assert not path.exists()
    for index in (0, 10000):
        rc, msg = oracle.write_cube(str(tmp_path / "o.cube"), grid, atoms, np.ones(8), orbital_index=index)
        assert rc == 1 and msg == "invalid orbital index for Cube"
    rc, msg = oracle.write_cube(str(tmp_path / "missing" / "x.cube"), grid, atoms, np.ones(8))
    assert rc == 3 and msg == "cannot open cube output"


def test_oracle_samples_read_back_within_half_a_digit(tmp_path):
    rng = np.random.default_rng(7)
    counts = [3, 4, 7]
    n = 84
    v = 10.0 ** rng.uniform(-40, 5, n) * rng.choice([-1.0, 1.0], n)
    v[:4] = [0.0, -0.0, 1.0, 999999.5e-6]
    grid = {"origin": [0.1, -0.2, 0.3], "steps": np.diag([0.2, 0.25, 0.3]), "counts": counts}
    atoms = (np.array([8, 1]), np.array([8.0, 1.0]), np.array([[0.0, 0.0, 0.0], [0.0, 1.4, 1.1]]))
    path = str(tmp_path / "r.cube")
    assert oracle.write_cube(path, grid, atoms, v)[0] == 0
    lines = open(path).read().split("\n")
    back = np.array(" ".join(lines[8:]).split(), dtype=np.float64)
    assert back.size == n
    assert np.all(np.abs(back - v) <= 5e-6 * np.abs(v) * (1 + 1e-12))
    # Lines: breaks after every sixth sample and after every run of 7.
    body = [ln for ln in lines[8:] if ln]
    breaks = sorted({i for i in range(1, n + 1) if i % 6 == 0 or i % 7 == 0})
    assert [len(ln) // 13 for ln in body] == list(np.diff([0] + breaks))
