"""Parity at the benchmark workloads' real shapes (synthetic molecules written
and read as FCHK): configuration 0 on its whole grid, slabs of configuration
1 in every contraction precision, the configuration 3 electrostatic
potential on the van der Waals shell, and size-independent invariants
(total = alpha + beta, Laplacian = trace, ELF in [0, 1]) over the whole
160^3 grid of configuration 1."""
import os

import numpy as np
import pytest

from kumo_hip import synthetic
from kumo_hip.api import Precision, PropertyKind as K
from support import oracle
from support.fchk import FchkData

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from kumo_hip.binding import Context
    c = Context(0)
    yield c
    c.close()


def load(ctx, fn, name):
    from kumo_hip.binding import Molecule
    path, mol, grid = synthetic.cached_fchk(fn, name)
    return Molecule(ctx, fchk_path=path), oracle.Molecule(FchkData(path)), grid


def test_config0_density_whole_grid(ctx):
    g, o, grid = load(ctx, lambda: synthetic.water_like(seed=0), "water_like_seed0")
    assert g.info["ao_mo_eligible"] == 1 and o.eligible
    got, used = g.evaluate([(K.Density, 0, 0, 0)], grid=grid)
    ref, rused = o.evaluate([(K.Density, 0, 0, 0)], grid=grid)
    assert used == rused == 2
    rho = ref[0][0]
    assert np.max(np.abs(got[0][0] - rho)) <= 1e-12 * rho.max()
    masks_g = g.masks(128, 0, grid=grid)
    masks_o = o.masks(128, 0, grid=grid)
    assert np.array_equal(masks_g, masks_o)
    g.close()


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32, Precision.BF16X3, Precision.BF16X6])
def test_config1_slab_precisions(ctx, precision):
    g, o, grid = load(ctx, lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
    props = [(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0), (K.Laplacian, 0, 0, 0), (K.Tau, 0, 0, 0),
             (K.Elf, 0, 0, 0), (K.Lol, 0, 0, 0)]
    # 40 screening blocks spread through the cube
    n = int(np.prod(grid["counts"]))
    blocks = np.linspace(0, n // 128 - 1, 40).astype(int)
    idx = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
    pts = oracle.grid_points(grid)[idx]
    got, _ = g.evaluate(props, points=pts, precision=precision)
    ref, _ = o.evaluate(props, points=pts)
    rho = ref[0][0]
    sig = rho >= 1e-6
    rel = np.max(np.abs(got[0][0] - rho)[sig] / rho[sig])
    elf = np.max(np.abs(got[4][0] - ref[4][0]))
    elf4 = np.max(np.abs(got[4][0] - ref[4][0])[rho >= 1e-4])
    print(f"[gpu] config1 precision {precision.name}: rho max rel {rel:.3g}, ELF max abs {elf:.3g} "
          f"(rho>=1e-4: {elf4:.3g})", flush=True)
    bound = {Precision.FP64: 1e-10, Precision.FP32: 1e-6, Precision.BF16X6: 1e-6, Precision.BF16X3: 1e-3}[precision]
    assert rel < bound
    assert np.all((got[4][0] >= 0) & (got[4][0] <= 1))
    g.close()


def test_config3_esp_on_vdw_shell(ctx):
    from kumo_hip.binding import Molecule
    mol = synthetic.pople_molecule()
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "kumo_pople20_seed3.fchk")
    mol.write_fchk(path)
    g = Molecule(ctx, fchk_path=path)
    o = oracle.Molecule(FchkData(path))
    pts = synthetic.vdw_shell_points(mol.atomic_numbers, mol.xyz, 3000)
    got, _ = g.evaluate([(K.ElectrostaticPotential, 0, 0, 0)], points=pts)
    ref = o.esp(pts)
    err = np.max(np.abs(got[0][0] - ref))
    print(f"[gpu] config3 ESP: {len(pts)} points, {o.esp_pair_count()} pairs, max abs diff {err:.3g}", flush=True)
    assert err <= 1e-6 * max(1.0, np.abs(ref).max())
    assert np.max(np.abs(got[0][0] - ref) / np.maximum(np.abs(ref), 1e-3)) < 1e-9
    g.close()


def test_config1_whole_grid_invariants(ctx):
    g, _, grid = load(ctx, lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
    props = [(K.Density, 0, 0, 0), (K.Density, 1, 0, 0), (K.Density, 2, 0, 0), (K.Laplacian, 0, 0, 0),
             (K.Hessian, 0, 0, 0), (K.Elf, 0, 0, 0)]
    got, _ = g.evaluate(props, grid=grid, precision=Precision.BF16X6)
    t, a, b = got[0][0], got[1][0], got[2][0]
    assert np.allclose(t, a + b, rtol=1e-12, atol=0)
    h = got[4]
    assert np.allclose(got[3][0], h[0] + h[3] + h[5], rtol=1e-12, atol=1e-300)
    assert np.all((got[5][0] >= 0) & (got[5][0] <= 1))
    assert np.all(t >= 0)
    g.close()
