"""Benchmark of the HIP FCHK->grid evaluator on one MI355X.

Workload (one step): a 40-atom molecule (16 C/N/O + 24 H, >= 1.8 bohr apart),
def2-TZVP-shaped pure shells (640 AOs), unrestricted 60 alpha / 59 beta
occupied MOs, evaluated on a 160^3 cube at 0.15 bohr (4 096 000 points):
total rho, grad rho, lap rho, tau, ELF and LOL, AUTO backend (AO-MO),
screening 1e-12, block 128. Molecule tables and MO coefficients are resident
in HBM and the outputs are written to HBM, so the timed region is the device
work of one evaluate() call per step.

Printed: one JSON line with points/s of the mixed-precision path (default
split-bf16 x6 with FP32 accumulation), the FP64 GPU path on the same
workload, the error of the mixed path against the FP64 oracle on a sample of
screening blocks, the roofline of the dominant contraction kernel (HIP events
on the library's stream) and the oracle (FP64 CPU restatement, OpenMP) timed
on a bounded sample of the same blocks.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "chem-kumo_amd"))

import numpy as np  # noqa: E402

from kumo_hip.api import Precision, PropertyKind as K, channels  # noqa: E402

PROPS = [(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0), (K.Laplacian, 0, 0, 0), (K.Tau, 0, 0, 0), (K.Elf, 0, 0, 0),
         (K.Lol, 0, 0, 0)]
PEAK = {"FP64": 78.6e12, "FP32": 157.3e12, "BF16X3": 2.5e15, "BF16X6": 2.5e15}
PRODUCTS = {"FP64": 1, "FP32": 1, "BF16X3": 3, "BF16X6": 6}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_gpu(mol, grid, precision, steps, warmup, torch):
    from kumo_hip.binding import Context  # noqa: F401
    n = int(np.prod(grid["counts"].astype(np.int64)))
    nch = sum(channels(p[0]) for p in PROPS)
    out = torch.empty((nch, n), dtype=torch.float64, device="cuda")
    ptrs = [out[c].data_ptr() for c in range(nch)]
    for _ in range(warmup):
        mol.evaluate_device(PROPS, grid, ptrs, precision=precision)
    torch.cuda.synchronize()
    kernel_ms, flops, fields_ms, fields_bytes = 0.0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        mol.evaluate_device(PROPS, grid, ptrs, precision=precision)
        st = mol.ctx.kernel_stats()
        kernel_ms += st["kernel_ms"]
        flops += st["flops"]
        fields_ms += st["fields_ms"]
        fields_bytes += st["fields_bytes"]
        launches = st["launches"]
        rescue_ms, rescue_points = st["rescue_ms"], st["rescue_points"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"seconds": dt, "points": n * steps, "kernel_ms": kernel_ms / steps, "flops": flops / steps,
            "launches": launches, "out": out, "rescue_ms": rescue_ms, "rescue_points": rescue_points,
            "fields_ms": fields_ms / steps, "fields_bytes": fields_bytes / steps}


CHANNELS = ["rho", "grad_x", "grad_y", "grad_z", "lap", "tau", "elf", "lol"]


OE_PROPS = [(K.AverageLocalIonizationEnergy, 0, 0, 0), (K.LocalElectronAffinity, 0, 0, 0),
            (K.LocalElectronAttachmentEnergy, 0, 0, 0), (K.LocalElectronegativity, 0, 0, 0),
            (K.LocalHardness, 0, 0, 0)]


def bench_orbital_energy(ctx, mol, grid, path, torch, no_cpu):
    """ALIE, LEA, LEAE, local electronegativity and hardness of the step's
    molecule on its cube (src/orbital_energy_fields.hpp): every MO of both
    channels contracted in FP64 on MFMA (weighted-sum epilogue), then the
    fields; the oracle on a sample of the points beside it."""
    n = int(np.prod(grid["counts"].astype(np.int64)))
    out = torch.empty((len(OE_PROPS), n), dtype=torch.float64, device="cuda")
    ptrs = [out[c].data_ptr() for c in range(len(OE_PROPS))]
    mol.evaluate_device(OE_PROPS, grid, ptrs)  # warm-up (operand table built on first use)
    torch.cuda.synchronize()
    t = time.perf_counter()
    mol.evaluate_device(OE_PROPS, grid, ptrs)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    st = ctx.kernel_stats()
    info = mol.info
    mos = info["nmo_alpha"] + (info["nmo_beta"] if info["unrestricted"] else 0)
    flops = 2.0 * info["nbasis"] * mos * n  # unscreened upper bound of the contraction
    res = {"points": n, "ms": dt * 1e3, "points_per_s": n / dt, "kernel_ms": st["kernel_ms"],
           "mos": mos, "precision": "FP64",
           "contraction": {"bound": "mfma", "unit": "TFLOP/s", "peak": 78.6,
                           "achieved_unscreened_upper": flops / (st["kernel_ms"] * 1e-3) / 1e12,
                           "note": "2 x AOs x MOs x points, every AO counted (the kept AOs are fewer)"}}
    log(f"orbital-energy fields: {res['points_per_s']:.4g} points/s ({res['ms']:.0f} ms, kernel "
        f"{st['kernel_ms']:.0f} ms, {mos} MOs)")
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from support import oracle
        from support.fchk import FchkData
        om = oracle.Molecule(FchkData(path))
        idx = np.arange(0, n, n // 2048)[:2048]
        pts = oracle.grid_points(grid)[idx]
        t = time.perf_counter()
        o, _ = om.evaluate(OE_PROPS, points=pts)
        cdt = time.perf_counter() - t
        got = out[:, torch.as_tensor(idx, device="cuda")].cpu().numpy()
        res["max_abs_error_vs_oracle"] = float(max(np.abs(got[i] - o[i][0]).max() for i in range(len(OE_PROPS))))
        res["cpu_baseline"] = {"value": len(idx) / cdt, "unit": "points/s", "cores": int(oracle.lib().orc_threads()),
                               "kind": "port", "sample": f"{len(idx)} points spread over the cube"}
        log(f"CPU oracle orbital-energy fields: {len(idx) / cdt:.4g} points/s")
    del out
    return res


def bench_surfaces(ctx, mol, path, no_cpu):
    """The ESP and ALIE surfaces of the step's molecule (rho = 0.001, 0.25
    bohr; src/surface.cpp): density grid, vertex placement on the GPU point
    path, the fields at the vertices, statistics. The oracle builds the same
    density isosurface beside it."""
    res = {}
    for name, call in (("esp", mol.electrostatic_potential_surface), ("alie", mol.local_ionization_energy_surface)):
        call()  # warm-up
        t = time.perf_counter()
        s = call()
        dt = time.perf_counter() - t
        res[name] = {"ms": dt * 1e3, "vertices": int(len(s.vertices)), "triangles": int(len(s.triangles)),
                     "average": s.statistics[0]["average"]}
        log(f"{name} surface: {dt * 1e3:.0f} ms, {len(s.vertices)} vertices")
    t = time.perf_counter()
    s = mol.density_isosurface()
    res["density_isosurface_ms"] = (time.perf_counter() - t) * 1e3
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from support import oracle
        from support.fchk import FchkData
        om = oracle.Molecule(FchkData(path))
        t = time.perf_counter()
        v, tr = oracle.density_isosurface(om)
        cdt = time.perf_counter() - t
        res["cpu_baseline"] = {"value": cdt * 1e3, "unit": "ms per density isosurface",
                               "cores": int(oracle.lib().orc_threads()), "kind": "port",
                               "sample": "the same density isosurface (grid densities, vertex placement)"}
        res["same_mesh_as_oracle"] = bool(np.array_equal(tr, s.triangles))
        log(f"CPU oracle density isosurface: {cdt * 1e3:.0f} ms (GPU {res['density_isosurface_ms']:.0f} ms)")
    return res


def bench_cube(ctx, mol, grid, out, no_cpu):
    """Cube files of the scalar fields of the step (rho, lap, tau, ELF, LOL),
    written from the device outputs (Field::write_cube, src/cube.cpp:70-111):
    wall time per file, device time of the text kernels and copies; the
    oracle's iostream writer (one thread) on a bounded sample beside it."""
    import shutil
    import tempfile
    atoms = mol.atoms()
    tmp = tempfile.mkdtemp(prefix="kumo_cube_")
    try:
        files = []
        for name, c in (("density", 0), ("laplacian", 4), ("tau", 5), ("elf", 6), ("lol", 7)):
            path = os.path.join(tmp, f"bench_{name}.cube")
            ctx.write_cube(path, grid, atoms, out[c])
            st = ctx.cube_stats()
            files.append(st)
            os.remove(path)
        n = files[0]["values"]
        wall = sum(f["total_ms"] for f in files) / len(files)
        kern = sum(f["kernel_ms"] for f in files) / len(files)
        copy = sum(f["copy_ms"] for f in files) / len(files)
        nbytes = sum(f["bytes"] for f in files) / len(files)
        hbm = 8.0 * n + nbytes  # samples read, text written
        res = {"files": len(files), "values_per_file": n, "bytes_per_file": nbytes, "ms_per_file": wall,
               "values_per_s": n / (wall * 1e-3), "text_kernel_ms": kern, "copy_ms": copy,
               "text_kernel": {"bound": "hbm", "achieved": hbm / (kern * 1e-3) / 1e9, "peak": 8000.0,
                               "unit": "GB/s", "frac": hbm / (kern * 1e-3) / 8e12,
                               "bytes": "8 B per sample read + the text written"}}
        log(f"Cube writer: {wall:.1f} ms per {n}-sample file ({nbytes / 1e6:.1f} MB), text kernels {kern:.2f} ms, "
            f"copies {copy:.2f} ms")
        if not no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            from support import oracle
            counts = np.array([64, 160, 160])
            m = int(np.prod(counts))
            sample = {"origin": grid["origin"], "steps": grid["steps"], "counts": counts}
            vals = out[0][:m].cpu().numpy()
            path = os.path.join(tmp, "cpu.cube")
            t = time.perf_counter()
            rc, msg = oracle.write_cube(path, sample, atoms, vals)
            dt = time.perf_counter() - t
            assert rc == 0, msg
            res["cpu_baseline"] = {"value": m / dt, "unit": "values/s", "cores": 1, "kind": "port",
                                   "sample": f"the first {m} density samples (64 x 160 x 160 slab) written by "
                                             "the oracle's iostream restatement of the writer"}
            res["speedup_vs_cpu"] = (n / (wall * 1e-3)) / (m / dt)
            log(f"CPU Cube writer: {m / dt:.4g} values/s")
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def full_grid_errors(got, ref, torch):
    """Per output channel: max abs error and max abs error relative to the
    channel's largest |value|, rho relative errors above density floors, and
    ELF errors where rho >= 1e-4, against the FP64 GPU path on every point."""
    out = {}
    rho = ref[0]
    for c, name in enumerate(CHANNELS):
        d = (got[c] - ref[c]).abs()
        out[name] = {"max_abs": float(d.max()), "max_abs_over_max": float(d.max() / ref[c].abs().max())}
    rel = (got[0] - rho).abs() / rho.clamp_min(1e-300)
    for floor in (1e-6, 1e-4, 1e-2):
        out["rho"][f"max_rel_rho_ge_{floor:g}"] = float(rel[rho >= floor].max())
    out["elf"]["max_abs_rho_ge_1e-4"] = float((got[6] - ref[6]).abs()[rho >= 1e-4].max())
    return out


PROPS_BATCH = [(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0), (K.Laplacian, 0, 0, 0), (K.Tau, 0, 0, 0),
               (K.Elf, 0, 0, 0)]


def bench_batch(args):
    """256 synthetic molecules (3-60 atoms, 6-31G*- or def2-TZVP-shaped, each
    on its own cube of edge 40-120 at 0.2 bohr), rho, grad, lap, tau, ELF,
    evaluated as one batch over 4 streams and replayed from a HIP graph."""
    import torch
    from kumo_hip import synthetic
    from kumo_hip.binding import Batch, Context, Molecule
    rng = np.random.default_rng(2)
    specs = [(int(rng.integers(3, 61)), "pople" if rng.random() < 0.5 else "tzvp", int(rng.integers(40, 121)))
             for _ in range(args.molecules)]
    ctx = Context(0)
    t = time.perf_counter()
    smols, grids = [], []
    for i, (na, fam, edge) in enumerate(specs):
        seed = 1000 + i
        while True:  # redraw the rare exponent sets that are numerically linearly dependent
            try:
                sm, grid = synthetic.batch_molecule(seed, na, fam, edge)
                break
            except ValueError:
                seed += 100000
        smols.append(sm)
        grids.append(grid)
    mols = [Molecule(ctx, data=sm) for sm in smols]
    counts = [int(np.prod(g["counts"].astype(np.int64))) for g in grids]
    total = sum(counts)
    log(f"{len(mols)} molecules ({sum(m.info['nbasis'] for m in mols)} AOs in all), {total} points, "
        f"ready in {time.perf_counter() - t:.1f}s")
    nch = sum(channels(p[0]) for p in PROPS_BATCH)
    res = {}
    outs = {}
    for prec in (args.precision, "FP64"):
        o = [torch.empty((nch, n), dtype=torch.float64, device="cuda") for n in counts]
        ptrs = [[t_[c].data_ptr() for c in range(nch)] for t_ in o]
        b = Batch(ctx, mols, grids, PROPS_BATCH, ptrs, streams=4, precision=int(Precision[prec]))
        for _ in range(max(1, args.warmup)):
            b.run(graph=True)
        ms = [b.run(graph=True) for _ in range(args.steps)]
        eager = b.run(graph=False)
        res[prec] = {"value": total / (min(ms) * 1e-3), "ms_per_step": min(ms), "ms_median": float(np.median(ms)),
                     "ms_without_graph": eager}
        log(f"{prec}: {res[prec]['value']:.4g} points/s, {min(ms):.1f} ms/step (graph), {eager:.1f} ms without")
        outs[prec] = o
        b.close()
    rel, elf = 0.0, 0.0
    for a, r in zip(outs[args.precision], outs["FP64"]):
        rho = r[0]
        sel = rho >= 1e-6
        if bool(sel.any()):
            rel = max(rel, float(((a[0] - rho).abs()[sel] / rho[sel]).max()))
        elf = max(elf, float((a[6] - r[6]).abs().max()))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from support import oracle
        pts_done, secs = 0, 0.0
        for i in range(0, len(mols), max(1, len(mols) // 16)):
            om = oracle.Molecule(smols[i])
            pts = oracle.grid_points(grids[i])
            sel = np.arange(0, len(pts) // 128, max(1, len(pts) // 128 // 32))[:32]
            idx = (sel[:, None] * 128 + np.arange(128)[None, :]).ravel()
            t0 = time.perf_counter()
            om.evaluate(PROPS_BATCH, points=pts[idx], block_size=128)
            secs += time.perf_counter() - t0
            pts_done += len(idx)
        cpu = {"value": pts_done / secs, "unit": "grid points/s", "cores": int(oracle.lib().orc_threads()),
               "kind": "port", "sample": f"32 screening blocks from each of 16 of the molecules ({pts_done} points)"}
        log(f"CPU oracle: {cpu['value']:.4g} points/s")
    out = {"metric": "grid pts/s for rho,grad rho,lap rho,tau,ELF vs FP64 GPU path at stated error (batched)",
           "value": res[args.precision]["value"], "unit": "grid points/s", "n_gpus": 1,
           "ms_per_step": res[args.precision]["ms_per_step"], "higher_is_better": True,
           "dtype": args.precision.lower(), "data": "synthetic (seeded molecules built in memory)",
           "config": {"workload": f"{len(mols)} molecules of 3-60 atoms (6-31G*/def2-TZVP shapes 50/50), cubes of "
                                  "edge 40-120 at 0.2 bohr, rho/grad/lap/tau/ELF, one graph replay per step",
                      "points": total, "streams": 4},
           "fp64_gpu": res["FP64"], "speedup_vs_fp64_gpu": res[args.precision]["value"] / res["FP64"]["value"],
           "errors_vs_fp64_gpu": {"rho_max_rel_rho_ge_1e-6": rel, "elf_max_abs": elf},
           "cpu_baseline": cpu}
    print(json.dumps(out))
    for m in mols:
        m.close()
    ctx.close()


def bench_esp(args):
    """Exact electrostatic potential of a 20-atom 6-31G*-shaped molecule on
    100 000 points of its 1.4 x van der Waals surface (FP64, Boys function)."""
    import torch
    from kumo_hip import synthetic
    from kumo_hip.binding import Context, Molecule
    sm = synthetic.pople_molecule()
    pts = synthetic.vdw_shell_points(sm.atomic_numbers, sm.xyz, 100000)
    ctx = Context(0)
    mol = Molecule(ctx, data=sm)
    prop = [(K.ElectrostaticPotential, 0, 0, 0)]
    ms = []
    for r in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        v, _ = mol.evaluate(prop, points=pts)
        if r >= args.warmup:
            ms.append(1e3 * (time.perf_counter() - t0))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from support import oracle
    om = oracle.Molecule(sm)
    sample = pts[:: len(pts) // 500][:500]
    t0 = time.perf_counter()
    ref = om.esp(sample)
    cpu_s = time.perf_counter() - t0
    got = v[0][0][:: len(pts) // 500][:500]
    pairs = om.esp_pair_count()
    out = {"metric": "ESP points/s (exact, FP64)", "value": len(pts) / (min(ms) * 1e-3), "unit": "points/s",
           "ms_per_step": min(ms), "pairs": pairs, "pair_points_per_s": pairs * len(pts) / (min(ms) * 1e-3),
           "max_abs_err_vs_oracle": float(np.max(np.abs(got - ref))),
           "max_rel_err_vs_oracle": float(np.max(np.abs(got - ref) / np.abs(ref))),
           "config": {"workload": f"20-atom 6-31G*-shaped molecule ({sm.nbasis} AOs), {len(pts)} points on the "
                                  "1.4 x vdW surface", "includes": "host->device points and device->host values"},
           "cpu_baseline": {"value": len(sample) / cpu_s, "unit": "points/s", "cores": int(oracle.lib().orc_threads()),
                            "kind": "port", "sample": f"{len(sample)} of the points"}}
    print(json.dumps(out))
    mol.close()
    ctx.close()


def bench_serving(args):
    """Serving load: resident synthetic molecules of 10-150 atoms
    (def2-TZVP-shaped), Poisson arrivals of point batches (40 %), lines of 512
    points (30 %), 256 x 256 planes (20 %) and cubes of edge 64-128 (10 %),
    properties drawn from rho, lap rho, ELF, RDG + sign(lambda2) rho and an MO
    amplitude; p50 / p99 latency per request kind and aggregate points/s."""
    from kumo_hip import synthetic
    from kumo_hip.binding import Context, Molecule, Server
    rng = np.random.default_rng(4)
    ctx = Context(0)
    t = time.perf_counter()
    smols = []
    for i in range(args.molecules):
        na, seed = int(rng.integers(10, 151)), 2000 + i
        while True:  # redraw the rare exponent sets that are numerically linearly dependent
            try:
                smols.append(synthetic.batch_molecule(seed, na, "tzvp", 40)[0])
                break
            except ValueError:
                seed += 100000
    mols = [Molecule(ctx, data=sm) for sm in smols]
    log(f"{len(mols)} resident molecules ({sum(m.info['nbasis'] for m in mols)} AOs in all) "
        f"ready in {time.perf_counter() - t:.1f}s")
    props_sets = [[(K.Density, 0, 0, 0)], [(K.Laplacian, 0, 0, 0)], [(K.Elf, 0, 0, 0)],
                  [(K.ReducedGradient, 0, 0, 0), (K.SignLambda2Density, 0, 0, 0)], None]
    srv = Server(ctx, workers=args.workers, chunk_points=args.chunk)
    # warm every molecule's lazily built tables and the workers' workspaces
    for m, sm in zip(mols, smols):
        srv.submit(m, [(K.Elf, 0, 0, 0), (K.SignLambda2Density, 0, 0, 0)], points=sm.xyz,
                   precision=int(Precision.BF16X6)).wait()
    results = []
    for rate in args.rates:
        n_req = int(rate * args.seconds)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
        reqs = []
        for _ in range(n_req):
            i = int(rng.integers(len(mols)))
            sm = smols[i]
            c = sm.xyz.mean(0)
            u = rng.random()
            if u < 0.4:
                kind = "points"
                where = {"points": c + rng.normal(size=(int(rng.integers(1, 1001)), 3)) * 3.0}
            elif u < 0.7:
                kind = "line"
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                where = {"points": c + np.linspace(-8, 8, 512)[:, None] * d}
            elif u < 0.9:
                kind = "plane"
                ax = np.linspace(-8, 8, 256)
                gx, gy = np.meshgrid(ax, ax, indexing="ij")
                where = {"points": np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], 1) + c}
            else:
                kind = "cube"
                e = int(rng.integers(64, 129))
                h = 16.0 / e
                where = {"grid": {"origin": c - 8.0, "steps": np.eye(3) * h, "counts": np.array([e, e, e])}}
            ps = props_sets[int(rng.integers(len(props_sets)))]
            if ps is None:
                ps = [(K.Orbital, 0, 0, int(sm.nalpha))]
            reqs.append((i, kind, ps, where))
        t0 = time.perf_counter()
        handles = []
        for (i, kind, ps, where), ta in zip(reqs, arrivals):
            delay = t0 + ta - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            handles.append((kind, srv.submit(mols[i], ps, precision=int(Precision.BF16X6), **where),
                            len(where["points"]) if "points" in where else int(np.prod(where["grid"]["counts"]))))
        lat = {}
        npts = 0
        for kind, h, n in handles:
            _, info = h.wait()
            lat.setdefault(kind, []).append(info["latency_ms"])
            npts += n
        wall = time.perf_counter() - t0
        allv = [x for v in lat.values() for x in v]
        entry = {"rate_per_s": rate, "requests": n_req, "seconds": wall, "points_per_s": npts / wall,
                 "p50_ms": float(np.percentile(allv, 50)), "p99_ms": float(np.percentile(allv, 99))}
        for kind, v in lat.items():
            entry[kind] = {"n": len(v), "p50_ms": float(np.percentile(v, 50)), "p99_ms": float(np.percentile(v, 99))}
        small = lat.get("points", []) + lat.get("line", [])
        entry["point_line_p99_ms"] = float(np.percentile(small, 99)) if small else None
        log(f"rate {rate}/s: p50 {entry['p50_ms']:.1f} ms, p99 {entry['p99_ms']:.1f} ms, point/line p99 "
            f"{entry['point_line_p99_ms']:.1f} ms, {entry['points_per_s']:.4g} points/s")
        results.append(entry)
    out = {"metric": "serving latency p50/p99 and aggregate points/s", "unit": "ms",
           "config": {"workload": f"{len(mols)} resident def2-TZVP-shaped molecules of 10-150 atoms; Poisson "
                                  "arrivals 40 % points (1-1000), 30 % lines (512), 20 % planes (256^2), 10 % cubes "
                                  "(64^3-128^3); rho, lap, ELF, RDG+sign(l2)rho or MO amplitude; BF16X6",
                      "workers": args.workers, "chunk_points": args.chunk},
           "loads": results}
    print(json.dumps(out))
    srv.close()
    for m in mols:
        m.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1, choices=[1, 2, 3, 4],
                    help="1: 160^3 cube (headline line); 2: batched molecules; 3: ESP; 4: serving load")
    ap.add_argument("--molecules", type=int, default=None, help="molecules of configurations 2 and 4")
    ap.add_argument("--rates", type=float, nargs="*", default=[20.0, 50.0], help="arrival rates (config 4)")
    ap.add_argument("--seconds", type=float, default=10.0, help="arrival window per rate (config 4)")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--chunk", type=int, default=131072)
    ap.add_argument("--precision", default="BF16X6", choices=list(PEAK))
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-blocks", type=int, default=0,
                    help="screening blocks in the CPU sample (0: every block of the grid)")
    ap.add_argument("--no-cube", action="store_true", help="skip the Cube-writer timing")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the orbital-energy field and surface timings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path would shard "
              "screening blocks across GPUs, each writing its slice of the field, with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.config == 2:
        args.molecules = args.molecules or 256
        return bench_batch(args)
    if args.config == 3:
        return bench_esp(args)
    if args.config == 4:
        args.molecules = args.molecules or 32
        return bench_serving(args)

    import torch
    from kumo_hip import synthetic
    from kumo_hip.binding import Context, Molecule

    t = time.perf_counter()
    path, smol, grid = synthetic.cached_fchk(lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
    log(f"synthetic molecule {smol.nbasis} AOs ready in {time.perf_counter() - t:.1f}s")
    ctx = Context(0)
    mol = Molecule(ctx, fchk_path=path)
    assert mol.info["ao_mo_eligible"] == 1, "synthetic MOs must reproduce P"
    n = int(np.prod(grid["counts"].astype(np.int64)))

    log(f"timing {args.precision}: warmup {args.warmup}, steps {args.steps}")
    main_run = run_gpu(mol, grid, int(Precision[args.precision]), args.steps, args.warmup, torch)
    value = main_run["points"] / main_run["seconds"]
    ms_per_step = 1e3 * main_run["seconds"] / args.steps
    log(f"{args.precision}: {value:.4g} points/s, {ms_per_step:.1f} ms/step, kernel {main_run['kernel_ms']:.1f} ms")

    fp64 = None
    if not args.no_fp64 and args.precision != "FP64":
        r = run_gpu(mol, grid, int(Precision.FP64), 1, 1, torch)
        fp64 = {"value": r["points"] / r["seconds"], "ms_per_step": 1e3 * r["seconds"], "kernel_ms": r["kernel_ms"],
                "tflops": r["flops"] / (r["kernel_ms"] * 1e-3) / 1e12 if r["kernel_ms"] > 0 else None}
        log(f"FP64 GPU path: {fp64['value']:.4g} points/s")
        ref_out = r["out"]
        del r
        full_errors = full_grid_errors(main_run["out"], ref_out, torch)
        log(f"full-grid errors vs FP64 GPU: rho rel (>=1e-6) {full_errors['rho']['max_rel_rho_ge_1e-06']:.3g}, "
            f"ELF abs {full_errors['elf']['max_abs']:.3g}")
    else:
        ref_out, full_errors = None, None

    kernel_s = main_run["kernel_ms"] * 1e-3
    achieved = main_run["flops"] / kernel_s if kernel_s > 0 else 0.0
    traffic = None  # HBM bytes per launch, from the FETCH_SIZE / WRITE_SIZE counter runs of this workload
    tpath = os.path.join(ROOT, "profiles", "traffic_aomo_bf16x6_tzvp640.json")
    if args.precision == "BF16X6" and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh)["traffic_bytes_per_launch"]
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK[args.precision] / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK[args.precision], "traffic": traffic,
                "kernel": "aomo_kernel", "kernel_ms": main_run["kernel_ms"],
                "launches_per_step": main_run["launches"],
                "algorithmic_flops_per_step": main_run["flops"],
                "mfma_products_per_flop": PRODUCTS[args.precision]}

    errors, cpu = None, None
    nblocks = n // 128
    want = args.cpu_blocks if args.cpu_blocks > 0 else nblocks
    blocks = np.arange(0, nblocks, max(1, nblocks // want))[:want]
    sample_idx = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from support import oracle
        from support.fchk import FchkData
        om = oracle.Molecule(FchkData(path))
        all_pts = oracle.grid_points(grid)
        pts = all_pts[sample_idx]
        whole = len(blocks) == nblocks
        t = time.perf_counter()
        o, _ = om.evaluate(PROPS, points=pts, block_size=128)
        dt = time.perf_counter() - t
        cpu = {"value": len(pts) / dt, "unit": "grid points/s", "cores": int(oracle.lib().orc_threads()),
               "kind": "port",
               "sample": ("every point of the 160^3 grid" if whole else
                          f"{len(blocks)} screening blocks of 128 points spread over the 160^3 grid ({len(pts)} "
                          "points)") + ", same molecule and properties, FP64, 128-point blocks"}
        # the same restatement on one host thread, over 256 blocks spread over the grid
        threads = int(oracle.lib().orc_threads())
        oracle.lib().orc_set_threads(1)
        b1 = np.arange(0, nblocks, max(1, nblocks // 256))[:256]
        p1 = all_pts[(b1[:, None] * 128 + np.arange(128)[None, :]).ravel()]
        t = time.perf_counter()
        om.evaluate(PROPS, points=p1, block_size=128)
        cpu["single_thread"] = {"value": len(p1) / (time.perf_counter() - t), "unit": "grid points/s",
                                "sample": f"{len(b1)} screening blocks of 128 points spread over the grid"}
        oracle.lib().orc_set_threads(threads)
        log(f"CPU oracle: {cpu['value']:.4g} points/s on {cpu['cores']} threads, "
            f"{cpu['single_thread']['value']:.4g} on one")
        got = main_run["out"][:, torch.as_tensor(sample_idx, device="cuda")].cpu().numpy()
        rho = o[0][0]
        sig = rho >= 1e-6
        rel = float(np.max(np.abs(got[0] - rho)[sig] / rho[sig]))
        elf_all = float(np.max(np.abs(got[6] - o[4][0])))
        elf_sig = float(np.max(np.abs(got[6] - o[4][0])[rho >= 1e-4]))
        errors = {"rho_max_rel": rel, "elf_max_abs": elf_all, "elf_max_abs_rho_ge_1e-4": elf_sig,
                  "vs": "FP64 oracle on the CPU sample"}
        if ref_out is not None:
            g64 = ref_out[:, torch.as_tensor(sample_idx, device="cuda")].cpu().numpy()
            errors["fp64_gpu_rho_max_rel"] = float(np.max(np.abs(g64[0] - rho)[sig] / rho[sig]))
            d64 = np.abs(g64[6] - o[4][0])
            errors["fp64_gpu_elf_max_abs"] = float(d64.max())
            # ELF where rho is so small that tau - |grad rho|^2 / 8 rho is
            # rounding noise in FP64 itself: largest rho at which either GPU
            # path differs from the oracle by more than 1e-5
            off = (np.abs(got[6] - o[4][0]) > 1e-5) | (d64 > 1e-5)
            errors["elf_points_off_by_1e-5"] = int(off.sum())
            errors["elf_off_max_rho"] = float(rho[off].max()) if off.any() else None

    cube = None if args.no_cube else bench_cube(ctx, mol, grid, main_run["out"], args.no_cpu)
    extras = {} if args.no_extras else {
        "orbital_energy_fields": bench_orbital_energy(ctx, mol, grid, path, torch, args.no_cpu),
        "surfaces": bench_surfaces(ctx, mol, path, args.no_cpu)}

    result = {
        "metric": "grid pts/s for rho,grad rho,lap rho,tau,ELF vs FP64 GPU path at stated error; p50/p99 latency",
        "value": value, "unit": "grid points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16" if args.precision.startswith("BF16") else args.precision.lower(),
        "data": "synthetic (seeded 40-atom TZVP-shaped molecule written and read as FCHK)",
        "config": {"workload": "40-atom def2-TZVP-shaped UHF (640 AOs, 60a/59b) on a 160^3 cube at 0.15 bohr: "
                               "rho, grad, lap, tau, ELF, LOL",
                   "points": n, "precision": args.precision, "block_size": 128, "backend": "AUTO(AO-MO)",
                   "parallelism": "single GPU"},
        "fp64_gpu": fp64,
        "speedup_vs_fp64_gpu": (value / fp64["value"]) if fp64 else None,
        "errors": errors,
        "errors_full_grid": dict(full_errors or {}, vs="FP64 GPU path, all points") if full_errors else None,
        "elf_rescue": {"points": main_run["rescue_points"], "fraction": main_run["rescue_points"] / n,
                       "fp64_kernel_ms": main_run["rescue_ms"], "criteria": "density or ELF uncertain"},
        "roofline": roofline,
        "field_reductions": {
            "bound": "hbm", "unit": "GB/s", "peak": 8000.0, "kernel": "fields_kernel",
            "kernel_ms": main_run["fields_ms"], "bytes_per_step": main_run["fields_bytes"],
            "achieved": main_run["fields_bytes"] / (main_run["fields_ms"] * 1e-3) / 1e9 if main_run["fields_ms"] else None,
            "frac": main_run["fields_bytes"] / (main_run["fields_ms"] * 1e-3) / 8e12 if main_run["fields_ms"] else None,
            "bytes": "channel moments read (2 x 13 x 8 B per point) + output channels written (8 x 8 B per point)"},
        "cpu_baseline": cpu,
        "cube_writer": cube,
        **extras,
    }
    print(json.dumps(result))
    mol.close()
    ctx.close()


if __name__ == "__main__":
    main()
