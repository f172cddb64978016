"""Serving loop (kumo_hip_server_*): requests submitted together, admitted
under a small HBM budget and cut into small ranges, give the same results as
the same requests evaluated alone (bit for bit: every screening block is
evaluated by the same kernels whatever the mix)."""
import numpy as np
import pytest

from kumo_hip.api import Precision, PropertyKind as K
from support.reference_cases import FCHK

pytestmark = pytest.mark.gpu


def _requests(rng):
    line = np.linspace([-3.0, -1.0, -0.5], [3.0, 1.5, 0.5], 512)
    plane = np.stack(np.meshgrid(np.linspace(-3, 3, 40), np.linspace(-3, 3, 37), [0.3], indexing="ij"), -1).reshape(-1, 3)
    cube = {"origin": np.array([-3.0, -3.0, -3.0]), "steps": np.eye(3) * 0.3, "counts": np.array([21, 20, 19])}
    props_sets = [
        [(K.Density, 0, 0, 0)],
        [(K.Laplacian, 0, 0, 0)],
        [(K.Elf, 0, 0, 0)],
        [(K.ReducedGradient, 0, 0, 0), (K.SignLambda2Density, 0, 0, 0)],
        [(K.Orbital, 0, 0, 2)],
        [(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0), (K.Tau, 0, 0, 0), (K.Elf, 0, 0, 0)],
    ]
    reqs = []
    for i in range(24):
        name = ["water", "acrolein", "o2"][i % 3]
        props = props_sets[i % len(props_sets)]
        kind = i % 4
        prec = [Precision.FP64, Precision.BF16X6][(i // 4) % 2]
        if kind == 0:
            where = {"points": rng.normal(size=(int(rng.integers(1, 1000)), 3)) * 2.0}
        elif kind == 1:
            where = {"points": line}
        elif kind == 2:
            where = {"points": plane}
        else:
            where = {"grid": cube}
        reqs.append((name, props, where, prec))
    return reqs


def test_results_same_alone_and_in_the_mix():
    from kumo_hip.binding import Context, Molecule, Server
    ctx = Context(0)
    mols = {n: Molecule(ctx, fchk_path=FCHK[n]) for n in ("water", "acrolein", "o2")}
    reqs = _requests(np.random.default_rng(3))
    alone = [mols[n].evaluate(p, precision=prec, **w)[0] for n, p, w, prec in reqs]
    srv = Server(ctx, workers=3, hbm_budget=2 << 20, chunk_points=1024)
    handles = [srv.submit(mols[n], p, precision=prec, **w) for n, p, w, prec in reqs]
    lat = []
    for (n, p, w, prec), h, ref in zip(reqs, handles, alone):
        got, info = h.wait()
        assert info["status"] == 0
        lat.append(info["latency_ms"])
        for a, b in zip(got, ref):
            assert np.array_equal(np.asarray(a), np.asarray(b)), (n, p)
    print(f"[gpu] server: {len(reqs)} requests, latency ms p50 {np.percentile(lat, 50):.2f} "
          f"max {max(lat):.2f}", flush=True)
    srv.close()
    for m in mols.values():
        m.close()
    ctx.close()


def test_invalid_request_is_rejected_at_submission():
    from kumo_hip.binding import Context, Molecule, Server
    ctx = Context(0)
    m = Molecule(ctx, fchk_path=FCHK["water"])
    srv = Server(ctx, workers=2)
    with pytest.raises(Exception):
        srv.submit(m, [(K.Orbital, 0, 0, 10_000)], points=np.zeros((4, 3)))
    got, info = srv.submit(m, [(K.Density, 0, 0, 0)], points=np.zeros((4, 3))).wait()
    assert info["status"] == 0 and got[0].shape == (1, 4)
    srv.close()
    m.close()
    ctx.close()
