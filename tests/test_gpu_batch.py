"""Batched evaluation of several molecules and grids (kumo_hip_batch_*):
every item equals its own kumo_hip_evaluate_grid result bit for bit, with
and without HIP graph replay, over several streams."""
import numpy as np
import pytest

from kumo_hip.api import Precision, PropertyKind as K, channels
from support.reference_cases import FCHK

pytestmark = pytest.mark.gpu

PROPS = [(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0), (K.Laplacian, 0, 0, 0), (K.Tau, 0, 0, 0), (K.Elf, 0, 0, 0)]


def _grid(origin, counts, step):
    return {"origin": np.asarray(origin, float), "steps": np.eye(3) * step, "counts": np.asarray(counts)}


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.BF16X6])
def test_batch_matches_single_evaluations(precision):
    import torch
    from kumo_hip.binding import Batch, Context, Molecule
    ctx = Context(0)
    names = ["water", "acrolein", "o2", "acrolein", "water"]
    mols = {n: Molecule(ctx, fchk_path=FCHK[n]) for n in set(names)}
    grids = [_grid([-4.0, -3.6, -3.2], [33, 30, 27], 0.25), _grid([-5.0, -4.0, -3.0], [41, 35, 22], 0.3),
             _grid([-3.0, -3.0, -4.0], [20, 21, 40], 0.2), _grid([-2.0, -1.0, -1.5], [9, 7, 5], 0.5),
             _grid([-1.0, -1.0, -1.0], [1, 1, 3], 0.1)]
    nch = sum(channels(p[0]) for p in PROPS)
    outs, ptrs = [], []
    for g in grids:
        n = int(np.prod(g["counts"]))
        t = torch.full((nch, n), float("nan"), dtype=torch.float64, device="cuda")
        outs.append(t)
        ptrs.append([t[c].data_ptr() for c in range(nch)])
    b = Batch(ctx, [mols[n] for n in names], grids, PROPS, ptrs, streams=3, precision=precision)
    for graph in (False, True, True):
        for t in outs:
            t.fill_(float("nan"))
        ms = b.run(graph=graph)
        print(f"[gpu] batch precision {int(precision)} graph={graph}: {ms:.3f} ms", flush=True)
        for name, g, t in zip(names, grids, outs):
            single, _ = mols[name].evaluate(PROPS, grid=g, precision=precision)
            got = t.cpu().numpy()
            ref = np.concatenate([np.atleast_2d(x) for x in single])
            assert np.array_equal(got, ref), (name, graph)
    b.close()
    for m in mols.values():
        m.close()
    ctx.close()
