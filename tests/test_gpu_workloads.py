"""The benchmark's batched and serving workloads at their real shapes, against
the oracle (tests/evaluator_parallel_test.cpp:52-110 is the reference's model:
results independent of how the work is shared).

Batched: 24 molecules of the 256-molecule workload (seeded as the benchmark
draws them: both basis families, its largest def2-TZVP-shaped molecule, its
largest cube), one HIP-graph batch over four streams in split-bf16 and in
FP64; sampled screening blocks of every item against the oracle, and a few
items against their lone evaluation bit for bit.

Serving: resident def2-TZVP-shaped molecules of 10 to 150 atoms, requests of
every kind (point batches, lines, planes, cubes) submitted with Poisson-spaced
arrivals so that they overlap on the device; every result equals the same
request evaluated alone bit for bit and matches the oracle on sampled points."""
import os
import sys
import time

import numpy as np
import pytest

from kumo_hip.api import Precision, PropertyKind as K, channels
from support import oracle

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the workload definitions)

PROPS = bench.PROPS_BATCH  # rho, grad rho, lap rho, tau, ELF
ELF_FLOOR = bench.ELF_FLOOR


def _pick_batch_items(specs, n=24):
    idx = [max(range(len(specs)), key=lambda i: (specs[i][1] == "tzvp", specs[i][0])),  # largest TZVP molecule
           max(range(len(specs)), key=lambda i: specs[i][2])]                            # largest cube
    rng = np.random.default_rng(21)
    for i in rng.permutation(len(specs)):
        if len(idx) == n:
            break
        if int(i) not in idx:
            idx.append(int(i))
    return idx


@pytest.fixture(scope="module")
def batch_items():
    specs = bench.config2_specs(256)
    idx = _pick_batch_items(specs)
    built = bench.parallel_map(bench._config2_molecule, [(i, specs[i]) for i in idx], spawn=True)
    fams = {specs[i][1] for i in idx}
    assert fams == {"pople", "tzvp"} and max(specs[i][0] for i in idx) >= 55 and max(specs[i][2] for i in idx) >= 118
    return idx, specs, built


def _errors(got, ref, rho):
    sig = rho >= 1e-6
    e = {"rho": float(np.max(np.abs(got[0] - rho)[sig] / rho[sig])) if sig.any() else 0.0}
    for name, chans in (("grad", (1, 2, 3)), ("lap", (4,)), ("tau", (5,))):
        e[name] = max(float(np.max(np.abs(got[c] - ref[c])) / max(np.max(np.abs(ref[c])), 1e-300)) for c in chans)
    fl = rho >= ELF_FLOOR
    e["elf"] = float(np.max(np.abs(got[6] - ref[6])[fl])) if fl.any() else 0.0
    return e


@pytest.mark.parametrize("precision", [Precision.BF16X6, Precision.FP64])
def test_batched_workload_against_oracle(batch_items, precision):
    import torch
    from kumo_hip.binding import Batch, Context, Molecule
    idx, specs, built = batch_items
    ctx = Context(0)
    mols = [Molecule(ctx, data=sm) for sm, _ in built]
    grids = [g for _, g in built]
    nch = sum(channels(p[0]) for p in PROPS)
    outs = [torch.full((nch, int(np.prod(g["counts"]))), float("nan"), dtype=torch.float64, device="cuda")
            for g in grids]
    b = Batch(ctx, mols, grids, PROPS, [[t[c].data_ptr() for c in range(nch)] for t in outs], streams=4,
              precision=int(precision))
    b.run(graph=True)   # first run: workspaces and autotuned tiles, eager
    for t in outs:
        t.fill_(float("nan"))
    b.run(graph=True)   # recorded into a graph
    for t in outs:
        t.fill_(float("nan"))
    ms = b.run(graph=True)  # replayed
    print(f"[gpu] batched workload {precision.name}: {len(mols)} items, "
          f"{sum(t.shape[1] for t in outs)} points, {ms:.1f} ms", flush=True)
    bounds = ({"rho": 1e-6, "grad": 1e-5, "lap": 1e-5, "tau": 1e-5, "elf": 1e-5} if precision == Precision.BF16X6 else
              {"rho": 1e-10, "grad": 1e-10, "lap": 1e-10, "tau": 1e-10, "elf": 2e-6})
    worst = {k: 0.0 for k in bounds}
    rng = np.random.default_rng(5)
    for k, ((sm, grid), t, i) in enumerate(zip(built, outs, idx)):
        n = t.shape[1]
        nb = (n + 127) // 128
        blocks = np.unique(np.concatenate([[(n // 2) // 128], rng.integers(0, nb, 3)]))
        sel = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
        sel = sel[sel < n]
        ref, _ = oracle.Molecule(sm).evaluate(PROPS, points=oracle.grid_points(grid)[sel], block_size=128)
        ref = np.concatenate([np.atleast_2d(x) for x in ref])
        got = t[:, torch.as_tensor(sel, device="cuda")].cpu().numpy()
        assert not np.isnan(got).any(), i
        e = _errors(got, ref, ref[0])
        for key in worst:
            worst[key] = max(worst[key], e[key])
    print(f"[gpu] batched workload {precision.name} vs oracle: " + ", ".join(f"{k} {v:.3g}" for k, v in worst.items()),
          flush=True)
    for key, bound in bounds.items():
        assert worst[key] <= bound, (key, worst[key])
    # items equal their lone evaluations bit for bit
    for k in (0, 1, len(mols) - 1):
        single, _ = mols[k].evaluate(PROPS, grid=grids[k], precision=int(precision))
        assert np.array_equal(outs[k].cpu().numpy(), np.concatenate([np.atleast_2d(x) for x in single])), idx[k]
    b.close()
    for m in mols:
        m.close()
    ctx.close()


SERVE_ATOMS = [150, 10, 96, 37, 64, 18, 120, 45]


@pytest.fixture(scope="module")
def serve_molecules():
    return bench.parallel_map(bench._config4_molecule, list(enumerate(SERVE_ATOMS)), spawn=True)


def _serve_requests(rng, smols, count):
    props_sets = [[(K.Density, 0, 0, 0)], [(K.Laplacian, 0, 0, 0)], [(K.Elf, 0, 0, 0)],
                  [(K.ReducedGradient, 0, 0, 0), (K.SignLambda2Density, 0, 0, 0)], None]
    reqs = []
    for r in range(count):
        i = int(rng.integers(len(smols)))
        sm = smols[i]
        c = sm.xyz.mean(0)
        kind = ["points", "line", "plane", "cube"][r % 4]
        if kind == "points":
            where = {"points": c + rng.normal(size=(int(rng.integers(1, 1001)), 3)) * 3.0}
        elif kind == "line":
            d = rng.normal(size=3)
            where = {"points": c + np.linspace(-8, 8, 512)[:, None] * (d / np.linalg.norm(d))}
        elif kind == "plane":
            ax = np.linspace(-8, 8, 96)
            gx, gy = np.meshgrid(ax, ax, indexing="ij")
            where = {"points": np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], 1) + c}
        else:
            e = int(rng.integers(40, 65))
            where = {"grid": {"origin": c - 8.0, "steps": np.eye(3) * (16.0 / e), "counts": np.array([e, e, e])}}
        ps = props_sets[r % len(props_sets)]
        if ps is None:
            ps = [(K.Orbital, 0, 0, int(sm.nalpha))]
        reqs.append((i, kind, ps, where))
    return reqs


def test_serving_workload_against_oracle_and_alone(serve_molecules):
    from kumo_hip.binding import Context, Molecule, Server
    smols = serve_molecules
    ctx = Context(0)
    mols = [Molecule(ctx, data=sm) for sm in smols]
    assert max(m.info["nbasis"] for m in mols) > 2000
    rng = np.random.default_rng(17)
    reqs = _serve_requests(rng, smols, 40)
    prec = int(Precision.BF16X6)
    alone = [mols[i].evaluate(ps, precision=prec, **w)[0] for i, _, ps, w in reqs]
    srv = Server(ctx, workers=4, chunk_points=65536)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, len(reqs)))
    t0 = time.perf_counter()
    handles = []
    for (i, kind, ps, w), ta in zip(reqs, arrivals):
        delay = t0 + ta - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        handles.append(srv.submit(mols[i], ps, precision=prec, **w))
    lat = []
    for (i, kind, ps, w), h, ref in zip(reqs, handles, alone):
        got, info = h.wait()
        assert info["status"] == 0
        lat.append(info["latency_ms"])
        for a, b in zip(got, ref):
            assert np.array_equal(np.asarray(a), np.asarray(b)), (i, kind, ps)
    print(f"[gpu] serving workload: {len(reqs)} requests on {len(mols)} molecules, latency p50 "
          f"{np.percentile(lat, 50):.1f} ms, max {max(lat):.1f} ms", flush=True)
    srv.close()
    # the oracle on sampled points of every request
    oms = {}
    for (i, kind, ps, w), ref in zip(reqs, alone):
        if i not in oms:
            oms[i] = oracle.Molecule(smols[i])
        om = oms[i]
        if "points" in w:
            pts = np.asarray(w["points"])
        else:
            pts = oracle.grid_points(w["grid"])
        n = len(pts)
        sel = np.unique(np.linspace(0, n - 1, min(n, 256)).astype(int))
        # the point path evaluates screening blocks: compare whole 128-point blocks
        blocks = np.unique(sel // 128)
        sel = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
        sel = sel[sel < n]
        o, _ = om.evaluate(ps, points=pts[sel], block_size=128)
        for p, a, b in zip(ps, ref, o):
            a = np.asarray(a)[:, sel]
            if p[0] == K.Elf:
                rho = om.evaluate([(K.Density, 0, 0, 0)], points=pts[sel], block_size=128)[0][0][0]
                assert np.max(np.abs(a[0] - b[0])[rho >= ELF_FLOOR], initial=0.0) <= 1e-5, (i, kind)
            elif p[0] == K.SignLambda2Density:
                # |sign(lambda2) rho| = rho, here without the density rescue (no density requested)
                assert np.max(np.abs(np.abs(a[0]) - np.abs(b[0])) / np.maximum(np.abs(b[0]), 1e-6)) <= 1e-5
            elif p[0] == K.ReducedGradient:
                rho = om.evaluate([(K.Density, 0, 0, 0)], points=pts[sel], block_size=128)[0][0][0]
                ok = rho >= 1e-4
                assert np.max(np.abs(a[0] - b[0])[ok] / np.abs(b[0])[ok], initial=0.0) <= 1e-5, (i, kind)
            elif p[0] == K.Orbital:
                assert np.max(np.abs(a[0] - b[0])) <= 1e-6 * np.max(np.abs(b[0])), (i, kind)
            else:
                for ca, cb in zip(a, b):
                    scale = np.max(np.abs(cb))
                    assert np.max(np.abs(ca - cb)) <= 1e-5 * scale, (i, kind, p)
    for m in mols:
        m.close()
    ctx.close()
