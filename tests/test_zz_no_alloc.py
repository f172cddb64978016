"""The allocation-free entry point (kumo_hip_evaluate_points_noalloc) against
kumo_hip_evaluate_points on the same device points: identical results, on two
calls in a row (the second smaller, so it runs over what the first left in
the shared scratch), in FP64 and in split-bf16 with its FP64 rescue, with more
points than one batch of the scratch holds; unsupported requests rejected."""
import numpy as np
import pytest

from kumo_hip import synthetic
from kumo_hip.api import Backend, DensityComponent as DC, Precision, PropertyKind as K, channels

pytestmark = pytest.mark.gpu

PROPS = [(K.Density, DC.Total, 0, 0), (K.Density, DC.Alpha, 0, 0), (K.Gradient, DC.Total, 0, 0),
         (K.Laplacian, DC.Total, 0, 0), (K.Tau, DC.Beta, 0, 0), (K.Elf, DC.Total, 0, 0), (K.Lol, DC.Total, 0, 0)]


@pytest.fixture(scope="module")
def setup():
    import torch
    from kumo_hip.binding import Context, Molecule
    path, sm, grid = synthetic.cached_fchk(lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
    ctx = Context(0)
    mol = Molecule(ctx, fchk_path=path)
    yield torch, ctx, mol, sm
    mol.close()
    ctx.close()


def run_noalloc(torch, mol, pts, precision, backend=Backend.AUTO):
    n = len(pts)
    dp = torch.as_tensor(np.ascontiguousarray(pts), device="cuda")
    nch = sum(channels(p[0]) for p in PROPS)
    out = torch.full((nch, n), np.nan, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()  # the inputs are ready before any stream reads them
    used = mol.evaluate_points_noalloc(PROPS, dp.data_ptr(), n, [out[c].data_ptr() for c in range(nch)],
                                       stream=stream, precision=precision, backend=backend)
    torch.cuda.synchronize()
    return out.cpu().numpy(), used


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.BF16X6])
def test_noalloc_equals_allocating_entry(setup, precision):
    torch, ctx, mol, sm = setup
    rng = np.random.default_rng(int(precision))
    centre = sm.xyz.mean(0)
    # 300 000 points: more than one batch of the 64 MiB scratch for this 640-AO molecule
    for n in (300_000, 1_000):
        pts = centre + rng.normal(size=(n, 3)) * 4.0
        print(f"[gpu] noalloc {precision.name}, {n} points", flush=True)
        got, used = run_noalloc(torch, mol, pts, int(precision))
        ref, rused = mol.evaluate(PROPS, points=pts, precision=int(precision))
        ref = np.concatenate([r for r in ref])
        assert used == rused == int(Backend.AO_MO)
        assert not np.isnan(got).any()
        assert np.array_equal(got, ref), float(np.nanmax(np.abs(got - ref)))


def test_noalloc_ao_p_and_rejections(setup):
    torch, ctx, mol, sm = setup
    pts = sm.xyz.mean(0) + np.random.default_rng(5).normal(size=(5_000, 3)) * 3.0
    got, used = run_noalloc(torch, mol, pts, int(Precision.FP64), backend=Backend.AO_P)
    ref, _ = mol.evaluate(PROPS, points=pts, backend=Backend.AO_P)
    assert used == int(Backend.AO_P)
    assert np.array_equal(got, np.concatenate([r for r in ref]))
    dp = torch.as_tensor(pts, device="cuda")
    out = torch.zeros((1, len(pts)), dtype=torch.float64, device="cuda")
    for props in ([(K.ElectrostaticPotential, 0, 0, 0)], [(K.Orbital, 0, 0, 1)]):
        with pytest.raises(ValueError):
            mol.evaluate_points_noalloc(props, dp.data_ptr(), len(pts), [out[0].data_ptr()])
    with pytest.raises(ValueError):  # a threshold whose screening radii are not resident
        mol.evaluate_points_noalloc([(K.Density, 0, 0, 0)], dp.data_ptr(), len(pts), [out[0].data_ptr()],
                                    screening_threshold=3e-9)
