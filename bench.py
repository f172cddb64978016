"""Benchmark of the HIP FCHK->grid evaluator on one MI355X.

Headline workload (one step): 256 synthetic molecules of 3-60 atoms (40 %
C/N/O, the rest H, >= 1.8 bohr apart), 6-31G*- or def2-TZVP-shaped bases
chosen 50/50, each on its own cube of edge 40-120 points at 0.2 bohr
(about 1.66e8 points in all): total rho, grad rho, lap rho, tau and ELF,
AUTO backend (AO-MO), screening 1e-12, blocks of 128 points. Every item is
issued into one HIP graph (recorded once, over 4 streams, tile shapes picked
per molecule by the autotuner during the warm-up) and a step is one replay
of that graph. Molecule tables, MO coefficients and outputs are resident in
HBM, so the timed region is device work only.

Printed: one JSON line with points/s of the mixed-precision path (split-bf16
x6 MFMA with FP32 accumulation, FP64 rescue of uncertain points), the FP64
GPU path on the same workload, the error of every property against the FP64
oracle on sampled screening blocks of every molecule and against the FP64
GPU path on every point, the roofline of the dominant contraction kernel
(HIP events on the streams it runs on), the oracle timed on a bounded sample
on the host cores, and two sub-records: the single 640-AO molecule on a
160^3 cube (the per-precision comparison) and a serving load (p50/p99
latency). --config 1, 3 or 4 runs one of those workloads on its own.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "chem-kumo_amd"))

import numpy as np  # noqa: E402

from kumo_hip.api import Precision, PropertyKind as K, channels  # noqa: E402

PROPS = [(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0), (K.Laplacian, 0, 0, 0), (K.Tau, 0, 0, 0), (K.Elf, 0, 0, 0),
         (K.Lol, 0, 0, 0)]
PROPS_BATCH = [(K.Density, 0, 0, 0), (K.Gradient, 0, 0, 0), (K.Laplacian, 0, 0, 0), (K.Tau, 0, 0, 0),
               (K.Elf, 0, 0, 0)]
BATCH_CHANNELS = ["rho", "grad_x", "grad_y", "grad_z", "lap", "tau", "elf"]
PEAK = {"FP64": 78.6e12, "FP32": 157.3e12, "BF16X3": 2.5e15, "BF16X6": 2.5e15}
PRODUCTS = {"FP64": 1, "FP32": 1, "BF16X3": 3, "BF16X6": 6}
METRIC = "grid pts/s for rho,grad rho,lap rho,tau,ELF vs FP64 GPU path at stated error; p50/p99 latency"
# ELF is compared where rho >= this floor: below it the Pauli term
# tau - |grad rho|^2 / 8 rho is a rounding of tau even in FP64 (one AO
# dominates every MO of a channel, so it vanishes in exact arithmetic) and two
# FP64 formulations of the same algorithm disagree (by up to ~1 below
# rho ~ 1e-20, by less than 1e-6 above 1e-14).
ELF_FLOOR = 1e-14


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """CPU model and core counts of the host (lscpu), and the threads the
    oracle uses."""
    info = {"model": None, "sockets": None, "cores_per_socket": None, "threads_per_core": None,
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        f = {k.strip(): v.strip() for k, v in f.items()}
        info.update(model=f.get("Model name"), sockets=f.get("Socket(s)"),
                    cores_per_socket=f.get("Core(s) per socket"), threads_per_core=f.get("Thread(s) per core"))
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    return info


def _config2_molecule(arg):
    """One molecule of the batched workload (worker process: one BLAS thread)."""
    i, (na, fam, edge) = arg
    from threadpoolctl import threadpool_limits
    from kumo_hip import synthetic
    with threadpool_limits(1):
        seed = 1000 + i
        while True:  # redraw the rare exponent sets that are numerically linearly dependent
            try:
                return synthetic.batch_molecule(seed, na, fam, edge)
            except ValueError:
                seed += 100000


def _config4_molecule(arg):
    i, na = arg
    from threadpoolctl import threadpool_limits
    from kumo_hip import synthetic
    with threadpool_limits(1):
        seed = 2000 + i
        while True:
            try:
                return synthetic.batch_molecule(seed, na, "tzvp", 40)[0]
            except ValueError:
                seed += 100000


def parallel_map(fn, items, spawn=False):
    """Builds synthetic molecules in worker processes (at most 16, the host
    share of one GPU). Forked workers need a parent that has not touched the
    GPU yet; spawn=True starts fresh interpreters instead (for callers that
    already use the GPU)."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing
    workers = max(1, min(16, len(items), (os.cpu_count() or 4)))
    ctx = multiprocessing.get_context("spawn" if spawn else "fork")
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        return list(ex.map(fn, items, chunksize=1))


def config2_specs(n):
    rng = np.random.default_rng(2)
    return [(int(rng.integers(3, 61)), "pople" if rng.random() < 0.5 else "tzvp", int(rng.integers(40, 121)))
            for _ in range(n)]


def errors_vs_oracle(gets, refs, floor=ELF_FLOOR):
    """Max errors of the BATCH_CHANNELS outputs against the oracle over the
    sampled points of every item: rho relative where rho >= 1e-6, gradient,
    Laplacian and tau relative to the channel's largest |value| in the item,
    ELF absolute where rho >= floor (and >= 1e-4)."""
    e = {"rho_max_rel": 0.0, "grad_max_rel_to_max": 0.0, "lap_max_rel_to_max": 0.0, "tau_max_rel_to_max": 0.0,
         "elf_max_abs": 0.0, "elf_max_abs_rho_ge_1e-4": 0.0, "points": 0, "points_below_elf_floor": 0}
    for got, ref in zip(gets, refs):
        rho = ref[0]
        sig = rho >= 1e-6
        if sig.any():
            e["rho_max_rel"] = max(e["rho_max_rel"], float(np.max(np.abs(got[0] - rho)[sig] / rho[sig])))
        for key, chans in (("grad_max_rel_to_max", (1, 2, 3)), ("lap_max_rel_to_max", (4,)),
                           ("tau_max_rel_to_max", (5,))):
            for c in chans:
                scale = np.max(np.abs(ref[c]))
                if scale > 0:
                    e[key] = max(e[key], float(np.max(np.abs(got[c] - ref[c])) / scale))
        d = np.abs(got[6] - ref[6])
        if (rho >= floor).any():
            e["elf_max_abs"] = max(e["elf_max_abs"], float(d[rho >= floor].max()))
        if (rho >= 1e-4).any():
            e["elf_max_abs_rho_ge_1e-4"] = max(e["elf_max_abs_rho_ge_1e-4"], float(d[rho >= 1e-4].max()))
        e["points"] += int(rho.size)
        e["points_below_elf_floor"] += int((rho < floor).sum())
    return e


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
        rescue_ms, rescue_points, rescue_flagged = st["rescue_ms"], st["rescue_points"], st["rescue_flagged"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"seconds": dt, "points": n * steps, "kernel_ms": kernel_ms / steps, "flops": flops / steps,
            "launches": launches, "out": out, "rescue_ms": rescue_ms, "rescue_points": rescue_points,
            "rescue_flagged": rescue_flagged,
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


def roofline_record(st, precision, traffic=None):
    """Roofline of the dominant contraction kernel from a timed run's launch
    statistics: algorithmic FLOPs per launch / average launch duration."""
    kernel_s = st["kernel_ms"] * 1e-3
    achieved = st["flops"] / kernel_s if kernel_s > 0 else 0.0
    n = max(1, st["launches"])
    return {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK[precision] / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK[precision], "traffic": traffic,
            "kernel": "aomo_wave_kernel" if precision in ("BF16X3", "BF16X6") else "aomo_kernel",
            "launches": st["launches"], "avg_launch_ms": st["kernel_ms"] / n,
            "algorithmic_flops_per_launch": st["flops"] / n, "mfma_products_per_flop": PRODUCTS[precision],
            "flops": "2 x kept AOs x occupied MOs x panels x points, summed over screening blocks"}


def load_traffic(name):
    """HBM bytes per launch of the dominant kernel, from the FETCH_SIZE and
    WRITE_SIZE counter runs summarised under profiles/ (null if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("traffic_bytes_per_launch")


def bench_config2(args, built, torch):
    """The headline: the batched 256-molecule workload, one graph replay per
    step, mixed precision beside the FP64 GPU path."""
    from kumo_hip.binding import Batch, Context, Molecule
    smols = [b[0] for b in built]
    grids = [b[1] for b in built]
    ctx = Context(0)
    mols = [Molecule(ctx, data=sm) for sm in smols]
    counts = [int(np.prod(g["counts"].astype(np.int64))) for g in grids]
    total = sum(counts)
    nao = sum(m.info["nbasis"] for m in mols)
    log(f"config 2: {len(mols)} molecules ({nao} AOs in all), {total} points, streams {args.streams}")
    nch = sum(channels(p[0]) for p in PROPS_BATCH)
    res, outs, stats = {}, {}, {}
    for prec in (args.precision, "FP64"):
        main = prec == args.precision
        o = [torch.empty((nch, n), dtype=torch.float64, device="cuda") for n in counts]
        ptrs = [[t_[c].data_ptr() for c in range(nch)] for t_ in o]
        b = Batch(ctx, mols, grids, PROPS_BATCH, ptrs, streams=args.streams, precision=int(Precision[prec]),
                  tile_points=args.tile_points if prec != "FP64" else 0)
        t = time.perf_counter()
        for _ in range(max(1, args.warmup if main else 1)):
            b.run(graph=True)
        log(f"{prec}: warm-up (workspaces, autotuned tiles, graph) {time.perf_counter() - t:.1f}s")
        steps = args.steps if main else max(1, min(args.steps, args.fp64_steps))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run(graph=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        eager = b.run(graph=False)
        # per-kernel figures from a run with the streams chained (no two
        # launches overlap, so each launch's event interval is its own time)
        st = b.timed_run(serial=True)
        res[prec] = {"value": total * steps / dt, "ms_per_step": 1e3 * dt / steps, "steps": steps,
                     "ms_without_graph": eager, "ms_serial": st["ms"]}
        stats[prec] = st
        log(f"{prec}: {res[prec]['value']:.4g} points/s, {res[prec]['ms_per_step']:.1f} ms/step (graph), "
            f"{eager:.1f} ms without, {st['ms']:.1f} ms serial; contraction {st['kernel_ms']:.1f} ms in "
            f"{st['launches']} launches, rescue {st['rescue_ms']:.1f} ms ({st['rescue_flagged']} points flagged, "
            f"{st['rescue_points']} slots), fields {st['fields_ms']:.1f} ms")
        outs[prec] = o
        b.close()
    # every point against the FP64 GPU path
    vs64 = {"rho_max_rel_rho_ge_1e-6": 0.0, "elf_max_abs": 0.0, "elf_max_abs_all_points": 0.0,
            "grad_max_rel_to_max": 0.0, "lap_max_rel_to_max": 0.0, "tau_max_rel_to_max": 0.0}
    for a, r in zip(outs[args.precision], outs["FP64"]):
        rho = r[0]
        sel = rho >= 1e-6
        if bool(sel.any()):
            vs64["rho_max_rel_rho_ge_1e-6"] = max(vs64["rho_max_rel_rho_ge_1e-6"],
                                                  float(((a[0] - rho).abs()[sel] / rho[sel]).max()))
        d = (a[6] - r[6]).abs()
        vs64["elf_max_abs_all_points"] = max(vs64["elf_max_abs_all_points"], float(d.max()))
        fl = rho >= ELF_FLOOR
        if bool(fl.any()):
            vs64["elf_max_abs"] = max(vs64["elf_max_abs"], float(d[fl].max()))
        for key, chans in (("grad_max_rel_to_max", (1, 2, 3)), ("lap_max_rel_to_max", (4,)),
                           ("tau_max_rel_to_max", (5,))):
            for c in chans:
                m = float(r[c].abs().max())
                if m > 0:
                    vs64[key] = max(vs64[key], float((a[c] - r[c]).abs().max()) / m)
    vs64["vs"] = f"FP64 GPU path, every point; ELF where rho >= {ELF_FLOOR:g} (and at all points)"
    # sampled screening blocks of every molecule against the oracle
    errors, cpu = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from support import oracle
        rng = np.random.default_rng(3)
        gets, refs, secs, npts = [], [], 0.0, 0
        per = args.oracle_blocks
        for i, (sm, grid) in enumerate(zip(smols, grids)):
            nb = (counts[i] + 127) // 128
            centre = (counts[i] // 2) // 128
            blocks = np.unique(np.concatenate([[centre], rng.integers(0, nb, per - 1)]))
            idx = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
            idx = idx[idx < counts[i]]
            om = oracle.Molecule(sm)
            pts = oracle.grid_points(grid)[idx]
            t = time.perf_counter()
            o, _ = om.evaluate(PROPS_BATCH, points=pts, block_size=128)
            secs += time.perf_counter() - t
            npts += len(idx)
            refs.append(np.concatenate([np.atleast_2d(x) for x in o]))
            gets.append(outs[args.precision][i][:, torch.as_tensor(idx, device="cuda")].cpu().numpy())
        errors = errors_vs_oracle(gets, refs)
        errors["vs"] = (f"FP64 oracle at {per} screening blocks of 128 points of each of the {len(mols)} molecules "
                        f"(the centre block and random ones); ELF where rho >= {ELF_FLOOR:g}")
        threads = int(oracle.lib().orc_threads())
        cpu = {"value": npts / secs, "unit": "grid points/s", "cores": threads, "kind": "port",
               "sample": f"{npts} points: {per} screening blocks of 128 points of each of the {len(mols)} molecules, "
                         "same properties, FP64, OpenMP over blocks", "host": host_cpu()}
        log(f"oracle: {cpu['value']:.4g} points/s on {threads} threads; errors {errors}")
    st = stats[args.precision]
    traffic = load_traffic("traffic_aomo_wave_bf16x6_config2.json") if args.precision == "BF16X6" else None
    out = {
        "metric": METRIC, "value": res[args.precision]["value"], "unit": "grid points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res[args.precision]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16" if args.precision.startswith("BF16") else args.precision.lower(),
        "data": "synthetic (seeded molecules built in memory; random-exponent bases, S-orthonormal random MOs)",
        "config": {"workload": f"{len(mols)} molecules of 3-60 atoms (6-31G*/def2-TZVP shapes 50/50), cubes of edge "
                               "40-120 at 0.2 bohr; rho, grad rho, lap rho, tau, ELF; one HIP-graph replay per step",
                   "molecules": len(mols), "points": total, "aos": nao, "streams": args.streams,
                   "precision": args.precision, "block_size": 128, "backend": "AUTO(AO-MO)",
                   "parallelism": "single GPU"},
        "fp64_gpu": res["FP64"],
        "speedup_vs_fp64_gpu": res[args.precision]["value"] / res["FP64"]["value"],
        "errors": errors,
        "errors_vs_fp64_gpu": vs64,
        "rescue": {"flagged_points": st["rescue_flagged"], "fraction": st["rescue_flagged"] / total,
                   "tile_slots": st["rescue_points"], "fp64_kernel_ms": st["rescue_ms"],
                   "criteria": "density or ELF uncertain at the precision"},
        "kernel_ms_serial": {"contraction": st["kernel_ms"], "rescue": st["rescue_ms"], "fields": st["fields_ms"],
                             "run": st["ms"], "note": "one run with the streams chained (no overlapping launches), "
                                                      "every launch bracketed by HIP events on its stream"},
        "roofline": roofline_record(st, args.precision, traffic),
        "roofline_fp64_path": roofline_record(stats["FP64"], "FP64"),
        "field_reductions": {"bound": "hbm", "unit": "GB/s", "peak": 8000.0, "kernel_ms": st["fields_ms"],
                             "bytes": st["fields_bytes"],
                             "achieved": st["fields_bytes"] / (st["fields_ms"] * 1e-3) / 1e9 if st["fields_ms"] else None},

        "cpu_baseline": cpu,
    }
    for m in mols:
        m.close()
    del outs
    ctx.close()
    return out


def bench_esp(args):
    """Electrostatic potential of a 20-atom 6-31G*-shaped molecule on 100 000
    points of its 1.4 x van der Waals surface (FP64, Boys function): the exact
    pair sums, and the far field at tolerances 1e-6 (the workload's) and 1e-8
    (the reference's default), each against the oracle's exact sum on a
    sample; `value` is the method the library picks by itself at 1e-6. Beside
    it the 640-AO def2-TZVP-shaped molecule on an 80^3 grid over its box."""
    from kumo_hip import synthetic
    from kumo_hip.binding import Context, Molecule
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from support import oracle
    from support.fchk import FchkData
    prop = [(K.ElectrostaticPotential, 0, 0, 0)]

    def timed(mol, **kw):
        ms, v = [], None
        for r in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            v, _ = mol.evaluate(prop, **kw)
            if r >= args.warmup:
                ms.append(1e3 * (time.perf_counter() - t0))
        return v[0][0], min(ms)

    def methods(mol, om, where, n, sample, sample_pts):
        t0 = time.perf_counter()
        ref = om.esp(sample_pts)
        cpu_s = time.perf_counter() - t0
        res = {}
        for name, tol, method in (("exact", 0.0, 2), ("far_1e-6", 1e-6, 1), ("far_1e-8", 1e-8, 1),
                                  ("auto_1e-6", 1e-6, 0)):
            v, ms = timed(mol, potential_tolerance=tol, esp_far_field=method, **where)
            err = np.abs(v[sample] - ref)
            res[name] = {"points_per_s": n / (ms * 1e-3), "ms": ms, "max_abs_err_vs_oracle": float(err.max()),
                         "max_rel_err_vs_oracle": float(np.max(err / np.abs(ref)))}
            print(f"[bench] ESP {name}: {res[name]['points_per_s']:.3e} points/s, {ms:.1f} ms, max abs err "
                  f"{err.max():.2e}", file=sys.stderr, flush=True)
        return res, len(sample_pts) / cpu_s

    sm = synthetic.pople_molecule()
    pts = synthetic.vdw_shell_points(sm.atomic_numbers, sm.xyz, 100000)
    ctx = Context(0)
    mol = Molecule(ctx, data=sm)
    om = oracle.Molecule(sm)
    sample = np.arange(0, len(pts), len(pts) // 500)[:500]
    res, cpu_rate = methods(mol, om, {"points": pts}, len(pts), sample, pts[sample])
    pairs = om.esp_pair_count()
    best = res["auto_1e-6"]
    out = {"metric": "ESP points/s (FP64, tolerance 1e-6)", "value": best["points_per_s"], "unit": "points/s",
           "ms_per_step": best["ms"], "pairs": pairs, "methods": res,
           "exact_pair_points_per_s": pairs * len(pts) / (res["exact"]["ms"] * 1e-3),
           "config": {"workload": f"20-atom 6-31G*-shaped molecule ({sm.nbasis} AOs), {len(pts)} points on the "
                                  "1.4 x vdW surface", "includes": "host->device points and device->host values"},
           "cpu_baseline": {"value": cpu_rate, "unit": "points/s", "cores": int(oracle.lib().orc_threads()),
                            "kind": "port", "sample": f"{len(sample)} of the points, exact sum"}}
    mol.close()
    path, _, grid1 = synthetic.cached_fchk(lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
    big = Molecule(ctx, fchk_path=path)
    obig = oracle.Molecule(FchkData(path))
    grid = {"origin": np.asarray(grid1["origin"]), "steps": np.asarray(grid1["steps"]) * 2.0,
            "counts": np.asarray([80, 80, 80], dtype=np.int32)}
    n = 80 ** 3
    gpts = oracle.grid_points(grid)
    gs = np.arange(11, n, n // 200)[:200]
    res2, cpu2 = methods(big, obig, {"grid": grid}, n, gs, gpts[gs])
    out["tzvp640_grid"] = {"workload": f"40-atom def2-TZVP-shaped molecule ({big.info['nbasis']} AOs), 80^3 grid at "
                                       "0.3 bohr", "pairs": obig.esp_pair_count(), "methods": res2,
                           "cpu_baseline": {"value": cpu2, "unit": "points/s", "kind": "port",
                                            "sample": f"{len(gs)} grid points, exact sum"}}
    big.close()
    print(json.dumps(out))
    ctx.close()


def bench_config1(args, c1, torch, sub=True):
    """The 640-AO UHF molecule on a 160^3 cube: rho, grad, lap, tau, ELF and
    LOL in the mixed precision and in FP64, errors against the oracle on
    sampled blocks and against the FP64 GPU path on every point."""
    from kumo_hip.binding import Context, Molecule
    path, smol, grid = c1
    ctx = Context(0)
    mol = Molecule(ctx, fchk_path=path)
    assert mol.info["ao_mo_eligible"] == 1, "synthetic MOs must reproduce P"
    n = int(np.prod(grid["counts"].astype(np.int64)))
    steps = min(args.steps, 10) if sub else args.steps
    warmup = min(args.warmup, 2) if sub else args.warmup
    main_run = run_gpu(mol, grid, int(Precision[args.precision]), steps, warmup, torch)
    value = main_run["points"] / main_run["seconds"]
    log(f"config 1 {args.precision}: {value:.4g} points/s, {1e3 * main_run['seconds'] / steps:.1f} ms/step")
    r = run_gpu(mol, grid, int(Precision.FP64), 1, 1, torch)
    fp64 = {"value": r["points"] / r["seconds"], "ms_per_step": 1e3 * r["seconds"], "kernel_ms": r["kernel_ms"],
            "tflops": r["flops"] / (r["kernel_ms"] * 1e-3) / 1e12 if r["kernel_ms"] > 0 else None}
    log(f"config 1 FP64: {fp64['value']:.4g} points/s")
    ref_out = r["out"]
    del r
    full = full_grid_errors(main_run["out"], ref_out, torch)
    errors, cpu = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from support import oracle
        from support.fchk import FchkData
        om = oracle.Molecule(FchkData(path))
        nblocks = n // 128
        want = 256 if sub else (args.cpu_blocks or nblocks)
        blocks = np.arange(0, nblocks, max(1, nblocks // want))[:want]
        idx = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
        pts = oracle.grid_points(grid)[idx]
        t = time.perf_counter()
        o, _ = om.evaluate(PROPS, points=pts, block_size=128)
        dt = time.perf_counter() - t
        cpu = {"value": len(pts) / dt, "unit": "grid points/s", "cores": int(oracle.lib().orc_threads()),
               "kind": "port", "sample": f"{len(blocks)} screening blocks of 128 points spread over the 160^3 grid"}
        got = main_run["out"][:, torch.as_tensor(idx, device="cuda")].cpu().numpy()
        g64 = ref_out[:, torch.as_tensor(idx, device="cuda")].cpu().numpy()
        ref = np.concatenate([np.atleast_2d(x) for x in o])  # rho, gx, gy, gz, lap, tau, elf, lol
        errors = errors_vs_oracle([got[:7]], [ref[:7]])
        errors["lol_max_abs"] = float(np.max(np.abs(got[7] - ref[7])))
        errors["fp64_gpu"] = errors_vs_oracle([g64[:7]], [ref[:7]])
        rho = ref[0]
        off = np.abs(g64[6] - ref[6]) > 1e-5
        errors["fp64_gpu"]["elf_points_off_by_1e-5_all_rho"] = int(off.sum())
        errors["fp64_gpu"]["elf_off_max_rho"] = float(rho[off].max()) if off.any() else None
        errors["vs"] = f"FP64 oracle at {len(idx)} points ({len(blocks)} blocks); ELF where rho >= {ELF_FLOOR:g}"
    kernel_s = main_run["kernel_ms"] * 1e-3
    st = {"kernel_ms": main_run["kernel_ms"], "launches": main_run["launches"], "flops": main_run["flops"]}
    traffic = load_traffic("traffic_aomo_wave_bf16x6_tzvp640.json") if args.precision == "BF16X6" else None
    res = {"workload": "40-atom def2-TZVP-shaped UHF (640 AOs, 60a/59b) on a 160^3 cube at 0.15 bohr: rho, grad, "
                       "lap, tau, ELF, LOL", "points": n, "precision": args.precision, "value": value,
           "unit": "grid points/s", "ms_per_step": 1e3 * main_run["seconds"] / steps, "steps": steps,
           "fp64_gpu": fp64, "speedup_vs_fp64_gpu": value / fp64["value"], "errors": errors,
           "errors_full_grid": dict(full, vs="FP64 GPU path, all points"),
           "rescue": {"flagged_points": main_run["rescue_flagged"], "fraction": main_run["rescue_flagged"] / n,
                      "tile_slots": main_run["rescue_points"], "fp64_kernel_ms": main_run["rescue_ms"]},
           "roofline": roofline_record(st, args.precision, traffic) if kernel_s > 0 else None,
           "cpu_baseline": cpu}
    if sub:
        del main_run, ref_out
        mol.close()
        ctx.close()
        return res
    return res, ctx, mol, main_run


def bench_config4(args, smols, rate, seconds):
    """Serving load over resident molecules: Poisson arrivals of point
    batches (40 %), lines of 512 points (30 %), 256 x 256 planes (20 %) and
    cubes of edge 64-128 (10 %); properties drawn from rho, lap rho, ELF,
    RDG + sign(lambda2) rho and an MO amplitude; p50 / p99 latency per kind
    and aggregate points/s."""
    from kumo_hip.binding import Context, Molecule, Server
    rng = np.random.default_rng(4)
    ctx = Context(0)
    mols = [Molecule(ctx, data=sm) for sm in smols]
    log(f"config 4: {len(mols)} resident molecules ({sum(m.info['nbasis'] for m in mols)} AOs in all)")
    props_sets = [[(K.Density, 0, 0, 0)], [(K.Laplacian, 0, 0, 0)], [(K.Elf, 0, 0, 0)],
                  [(K.ReducedGradient, 0, 0, 0), (K.SignLambda2Density, 0, 0, 0)], None]
    srv = Server(ctx, workers=args.workers, chunk_points=args.chunk)
    # warm every molecule's lazily built tables and the workers' workspaces
    for m, sm in zip(mols, smols):
        srv.submit(m, [(K.Elf, 0, 0, 0), (K.SignLambda2Density, 0, 0, 0)], points=sm.xyz,
                   precision=int(Precision.BF16X6)).wait()
    results = []
    for r in ([rate] if np.isscalar(rate) else rate):
        n_req = int(r * seconds)
        arrivals = np.cumsum(rng.exponential(1.0 / r, n_req))
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
        entry = {"rate_per_s": r, "requests": n_req, "seconds": wall, "points_per_s": npts / wall,
                 "p50_ms": float(np.percentile(allv, 50)), "p99_ms": float(np.percentile(allv, 99))}
        for kind, v in lat.items():
            entry[kind] = {"n": len(v), "p50_ms": float(np.percentile(v, 50)), "p99_ms": float(np.percentile(v, 99))}
        small = lat.get("points", []) + lat.get("line", [])
        entry["point_line_p99_ms"] = float(np.percentile(small, 99)) if small else None
        log(f"config 4 rate {r}/s: p50 {entry['p50_ms']:.1f} ms, p99 {entry['p99_ms']:.1f} ms, point/line p99 "
            f"{entry['point_line_p99_ms']:.1f} ms, {entry['points_per_s']:.4g} points/s")
        results.append(entry)
    out = {"metric": "serving latency p50/p99 and aggregate points/s", "unit": "ms",
           "workload": f"{len(mols)} resident def2-TZVP-shaped molecules of 10-150 atoms; Poisson arrivals 40 % points "
                       "(1-1000), 30 % lines (512), 20 % planes (256^2), 10 % cubes (64^3-128^3); rho, lap, ELF, "
                       "RDG+sign(l2)rho or MO amplitude; BF16X6 with its FP64 rescues",
           "workers": args.workers, "chunk_points": args.chunk, "loads": results}
    srv.close()
    for m in mols:
        m.close()
    ctx.close()
    return out


def build_inputs(args, need):
    """Every synthetic molecule of the requested workloads, built in worker
    processes before this process touches the GPU."""
    from kumo_hip import synthetic
    t = time.perf_counter()
    tasks = []
    if "c2" in need:
        tasks += [("c2", a) for a in enumerate(config2_specs(args.molecules))]
    if "c4" in need:
        rng = np.random.default_rng(4)
        tasks += [("c4", (i, int(rng.integers(10, 151)))) for i in range(args.serve_molecules)]
    built = parallel_map(_build_task, tasks) if tasks else []
    out = {"c2": [b for (k, _), b in zip(tasks, built) if k == "c2"],
           "c4": [b for (k, _), b in zip(tasks, built) if k == "c4"]}
    if "c1" in need:
        out["c1"] = synthetic.cached_fchk(lambda: synthetic.tzvp_cube(seed=1), "tzvp40_seed1")
    log(f"synthetic inputs ready in {time.perf_counter() - t:.1f}s")
    return out


def _build_task(task):
    kind, arg = task
    return _config2_molecule(arg) if kind == "c2" else _config4_molecule(arg)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4],
                    help="2: the batched molecules (headline, with configuration 1 and 4 sub-records); "
                         "1: 160^3 cube; 3: ESP; 4: serving load")
    ap.add_argument("--molecules", type=int, default=256, help="molecules of configuration 2")
    ap.add_argument("--serve-molecules", type=int, default=32, help="resident molecules of configuration 4")
    ap.add_argument("--streams", type=int, default=4, help="streams of the configuration 2 graph")
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed steps of the FP64 comparison (config 2)")
    ap.add_argument("--oracle-blocks", type=int, default=4, help="oracle blocks per molecule (config 2)")
    ap.add_argument("--rates", type=float, nargs="*", default=[20.0, 50.0], help="arrival rates (config 4)")
    ap.add_argument("--seconds", type=float, default=10.0, help="arrival window per rate (config 4)")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--chunk", type=int, default=131072)
    ap.add_argument("--precision", default="BF16X6", choices=list(PEAK))
    ap.add_argument("--tile-points", type=int, default=0, choices=[0, 16, 32, 64],
                    help="config 2: AO-MO workgroup tile of the reduced precision (0: the library's choice)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle comparisons")
    ap.add_argument("--no-sub", action="store_true", help="config 2 without the configuration 1 and 4 sub-records")
    ap.add_argument("--cpu-blocks", type=int, default=0,
                    help="config 1 alone: screening blocks in the CPU sample (0: every block of the grid)")
    ap.add_argument("--extras", action="store_true",
                    help="config 1 alone: also the Cube writer, orbital-energy fields and surfaces")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device once (the calibration kernel of the HBM traffic counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path would shard "
              "screening blocks or molecules across GPUs, each writing its slice of the fields, with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.config == 3:
        return bench_esp(args)
    need = {2: {"c2"} | (set() if args.no_sub else {"c1", "c4"}), 1: {"c1"}, 4: {"c4"}}[args.config]
    inputs = build_inputs(args, need)
    import torch
    if args.config == 4:
        print(json.dumps(bench_config4(args, inputs["c4"], args.rates, args.seconds)))
        return
    if args.config == 1:
        res, ctx, mol, main_run = bench_config1(args, inputs["c1"], torch, sub=False)
        grid, path = inputs["c1"][2], inputs["c1"][0]
        if args.extras:
            res["cube_writer"] = bench_cube(ctx, mol, grid, main_run["out"], args.no_cpu)
            res["orbital_energy_fields"] = bench_orbital_energy(ctx, mol, grid, path, torch, args.no_cpu)
            res["surfaces"] = bench_surfaces(ctx, mol, path, args.no_cpu)
        out = {"metric": METRIC, "value": res["value"], "unit": "grid points/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None,
               "dtype": "bf16" if args.precision.startswith("BF16") else args.precision.lower(),
               "data": "synthetic (seeded 40-atom TZVP-shaped molecule written and read as FCHK)",
               "config": {"workload": res.pop("workload"), "points": res["points"], "precision": args.precision,
                          "block_size": 128, "backend": "AUTO(AO-MO)", "parallelism": "single GPU"}}
        out.update({k: v for k, v in res.items() if k not in ("value", "ms_per_step", "points")})
        print(json.dumps(out))
        mol.close()
        ctx.close()
        return
    out = bench_config2(args, inputs["c2"], torch)
    if not args.no_sub:
        out["config1"] = bench_config1(args, inputs["c1"], torch)
        out["config4"] = bench_config4(args, inputs["c4"], 30.0, 6.0)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
