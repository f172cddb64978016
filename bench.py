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
    kernel_ms, flops = 0.0, 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        mol.evaluate_device(PROPS, grid, ptrs, precision=precision)
        st = mol.ctx.kernel_stats()
        kernel_ms += st["kernel_ms"]
        flops += st["flops"]
        launches = st["launches"]
        rescue_ms, rescue_points = st["rescue_ms"], st["rescue_points"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"seconds": dt, "points": n * steps, "kernel_ms": kernel_ms / steps, "flops": flops / steps,
            "launches": launches, "out": out, "rescue_ms": rescue_ms, "rescue_points": rescue_points}


CHANNELS = ["rho", "grad_x", "grad_y", "grad_z", "lap", "tau", "elf", "lol"]


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="BF16X6", choices=list(PEAK))
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-blocks", type=int, default=1024, help="screening blocks in the CPU sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path would shard "
              "screening blocks across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

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
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK[args.precision] / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK[args.precision], "traffic": None,
                "kernel": "aomo_kernel", "kernel_ms": main_run["kernel_ms"],
                "launches_per_step": main_run["launches"],
                "algorithmic_flops_per_step": main_run["flops"],
                "mfma_products_per_flop": PRODUCTS[args.precision]}

    errors, cpu = None, None
    blocks = np.arange(0, n // 128, max(1, (n // 128) // args.cpu_blocks))[:args.cpu_blocks]
    sample_idx = (blocks[:, None] * 128 + np.arange(128)[None, :]).ravel()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from support import oracle
        from support.fchk import FchkData
        om = oracle.Molecule(FchkData(path))
        pts = oracle.grid_points(grid)[sample_idx]
        t = time.perf_counter()
        o, _ = om.evaluate(PROPS, points=pts, block_size=128)
        dt = time.perf_counter() - t
        cpu = {"value": len(pts) / dt, "unit": "grid points/s", "cores": int(oracle.lib().orc_threads()),
               "kind": "port", "sample": f"{len(blocks)} screening blocks of 128 points spread over the 160^3 grid "
                                         f"({len(pts)} points), same molecule and properties, FP64"}
        log(f"CPU oracle: {cpu['value']:.4g} points/s on {cpu['cores']} threads")
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
                       "fp64_kernel_ms": main_run["rescue_ms"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))
    mol.close()
    ctx.close()


if __name__ == "__main__":
    main()
