"""The GPU Cube writer through the C-ABI (kumo_hip_write_cube,
kumo_hip_write_orbital_cube) against the oracle's restatement of
Field::write_cube / write_orbital_cube (src/cube.cpp:70-111, 253-262): the
files are byte-identical, for host and device samples, grids whose third axis
is or is not a multiple of six, runs of one sample, and fields longer than one
formatting chunk (2^21 samples). The GPU formatter on its own
(kumo_hip_format_scientific) equals Python's printf-style formatting (correctly
rounded, ties to even, as glibc) on random bit patterns over the whole double
range, exact decimal ties and powers of ten, on its fast path and with every
rounding taken on its exact path. Errors follow src/cube.cpp:75-80, 258-259:
nonfinite samples and orbital indices 0 or above 9999 are invalid arguments
and leave an existing file untouched; an unopenable path is a runtime error.
"""
import os

import numpy as np
import pytest

from support import oracle
from support.reference_cases import FCHK

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from kumo_hip.binding import Context
    c = Context(0)
    yield c
    c.close()


def special_values(rng):
    v = [0.0, -0.0, 5e-324, 2.2250738585072009e-308, 2.2250738585072014e-308, 1.7976931348623157e308, 1.0, 0.5,
         2.5, 1234565.0, 1234575.0, 100000.5, 9999995.0, 9.999995, 999999.5e-6]
    for e in range(-323, 309):
        p = float(f"1e{e}")
        v += [p, np.nextafter(p, 0.0), np.nextafter(p, np.inf), 9.999995 * p]
    # Exact ties of six significant digits: odd multiples of 10^k / 2.
    for k in range(0, 10):
        for f in rng.integers(100000, 1000000, 40):
            t = float((2 * int(f) + 1) * 5 ** k) * 2.0 ** (k - 1)
            if (2 * int(f) + 1) * 5 ** k < 2 ** 53:
                v += [t, np.nextafter(t, 0.0), np.nextafter(t, np.inf)]
    v = np.array([x for x in v if np.isfinite(x)])
    return np.concatenate([v, -v])


def random_bits(rng, n):
    bits = rng.integers(0, 2 ** 63, n, dtype=np.int64).astype(np.uint64) | \
        (rng.integers(0, 2, n).astype(np.uint64) << np.uint64(63))
    v = bits.view(np.float64)
    return v[np.isfinite(v)]


@pytest.mark.parametrize("precision", [5, 12, 0])
def test_formatter_matches_printf(ctx, precision):
    rng = np.random.default_rng(precision + 11)
    v = np.concatenate([special_values(rng), random_bits(rng, 60000),
                        10.0 ** rng.uniform(-50, 10, 20000) * rng.choice([-1.0, 1.0], 20000)])
    width = precision + 8
    want = ["%*.*e" % (width, precision, x) for x in v]
    print(f"[gpu] format_scientific P={precision} n={v.size}", flush=True)
    got = ctx.format_scientific(v, precision)
    bad = [(x, g, w) for x, g, w in zip(v, got, want) if g != w]
    assert not bad, bad[:5]
    sub = v[::7]
    got = ctx.format_scientific(sub, precision, exact_only=True)
    bad = [(x, g, w) for x, g, w in zip(sub, got, want[::7]) if g != w]
    assert not bad, bad[:5]


def make_case(rng, counts, with_specials=True):
    n = int(np.prod(counts))
    v = 10.0 ** rng.uniform(-30, 3, n) * rng.choice([-1.0, 1.0], n)
    if with_specials:
        s = special_values(rng)
        m = min(n, s.size)
        v[rng.choice(n, m, replace=False)] = s[:m]
    steps = np.array([[0.21, 0.0, 0.01], [0.0, 0.19, -0.02], [-0.03, 0.0, 0.2]])
    grid = {"origin": [-3.5, 2.25, -0.0], "steps": steps, "counts": list(counts)}
    natoms = int(rng.integers(1, 5))
    z = rng.integers(1, 30, natoms)
    q = z.astype(float) - rng.integers(0, 2, natoms) * 10.0
    atoms = (z, q, rng.normal(size=(natoms, 3)) * 3.0)
    return grid, atoms, v


def read(path):
    with open(path, "rb") as fh:
        return fh.read()


@pytest.mark.parametrize("counts", [(2, 2, 2), (3, 4, 7), (2, 3, 1), (1, 1, 1), (5, 3, 12), (4, 5, 74), (3, 1, 300)])
def test_cube_bytes_match_oracle(ctx, tmp_path, counts):
    rng = np.random.default_rng(sum(counts))
    grid, atoms, v = make_case(rng, counts)
    print(f"[gpu] write_cube counts={counts}", flush=True)
    for orbital, comment in ((None, None), (None, "ELF of a test field"), (7, None), (9999, "MO 9999")):
        got, want = str(tmp_path / "g.cube"), str(tmp_path / "o.cube")
        ctx.write_cube(got, grid, atoms, v, comment=comment, orbital_index=orbital)
        rc, msg = oracle.write_cube(want, grid, atoms, v, comment=comment, orbital_index=orbital)
        assert rc == 0, msg
        assert read(got) == read(want), (counts, orbital)


@pytest.mark.parametrize("counts", [(2, 1031, 1019), (1, 1, (1 << 21) + 77)])
def test_cube_longer_than_one_chunk(ctx, tmp_path, counts):
    rng = np.random.default_rng(5)
    grid, atoms, v = make_case(rng, counts)
    print(f"[gpu] write_cube counts={counts}", flush=True)
    got, want = str(tmp_path / "g.cube"), str(tmp_path / "o.cube")
    ctx.write_cube(got, grid, atoms, v)
    assert oracle.write_cube(want, grid, atoms, v)[0] == 0
    a, b = read(got), read(want)
    assert len(a) == len(b)
    assert a == b
    st = ctx.cube_stats()
    assert st["values"] == v.size and st["bytes"] == len(a)


def test_device_samples(ctx, tmp_path):
    import torch
    rng = np.random.default_rng(3)
    grid, atoms, v = make_case(rng, (6, 7, 29))
    t = torch.from_numpy(v).to("cuda:0")
    torch.cuda.synchronize()
    print("[gpu] write_cube device samples", flush=True)
    ctx.write_cube(str(tmp_path / "d.cube"), grid, atoms, t)
    ctx.write_cube(str(tmp_path / "h.cube"), grid, atoms, v)
    assert read(str(tmp_path / "d.cube")) == read(str(tmp_path / "h.cube"))


def test_evaluated_density_cube(ctx, tmp_path):
    from kumo_hip.api import DensityComponent as DC, Property, PropertyKind as K
    from kumo_hip.binding import Molecule
    mol = Molecule(ctx, fchk_path=FCHK["o2"])
    try:
        grid = {"origin": [-3.0, -3.0, -4.0], "steps": np.eye(3) * 0.25, "counts": [25, 25, 33]}
        print("[gpu] evaluate o2 density for a cube", flush=True)
        outs, _ = mol.evaluate([Property(K.Density, DC.Total)], grid=grid)
        rho = outs[0][0]
        atoms = mol.atoms()
        got, want = str(tmp_path / "g.cube"), str(tmp_path / "o.cube")
        ctx.write_cube(got, grid, atoms, rho)
        assert oracle.write_cube(want, grid, atoms, rho)[0] == 0
        assert read(got) == read(want)
        text = read(got).decode().split("\n")
        assert text[0] == "chem-kumo 2.1.0" and text[1] == "Electron density"
        back = np.array(" ".join(text[6 + atoms[0].size:]).split(), dtype=np.float64)
        assert np.all(np.abs(back - rho) <= 5e-6 * np.abs(rho) * (1 + 1e-12))
    finally:
        mol.close()


def test_rejections_leave_files_untouched(ctx, tmp_path):
    rng = np.random.default_rng(9)
    grid, atoms, v = make_case(rng, (2, 2, 3), with_specials=False)
    path = tmp_path / "keep.cube"
    path.write_bytes(b"previous contents\n")
    for bad in (np.nan, np.inf, -np.inf):
        w = v.copy()
        w[4] = bad
        with pytest.raises(ValueError, match="nonfinite Cube sample"):
            ctx.write_cube(str(path), grid, atoms, w)
        assert path.read_bytes() == b"previous contents\n"
    for index in (0, 10000):
        with pytest.raises(ValueError, match="invalid orbital index for Cube"):
            ctx.write_cube(str(path), grid, atoms, v, orbital_index=index)
        assert path.read_bytes() == b"previous contents\n"
    with pytest.raises(RuntimeError, match="cannot open cube output"):
        ctx.write_cube(str(tmp_path / "no" / "such" / "dir.cube"), grid, atoms, v)
    with pytest.raises(RuntimeError, match="cannot open orbital cube output"):
        ctx.write_cube(str(tmp_path / "no" / "dir.cube"), grid, atoms, v, orbital_index=3)
    bad_grid = dict(grid, counts=[2, 0, 3])
    with pytest.raises(ValueError):
        ctx.write_cube(str(path), bad_grid, atoms, v)
    # The context still works after the errors.
    ctx.write_cube(str(path), grid, atoms, v)
    assert oracle.write_cube(str(tmp_path / "o.cube"), grid, atoms, v)[0] == 0
    assert path.read_bytes() == read(str(tmp_path / "o.cube"))
    assert os.path.getsize(path) > 0
