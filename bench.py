"""Benchmark of the crabSAXS profile hot path on one MI355X.

Workload (one "step"): the Debye precision-study profile -- a seeded 20,000-atom
globule (R = 35 A), q = linspace(0, 1, 500), default solvent parameters --
computed with the FP32-compensated per-type pair-table kernel and contracted to
I(q).  Inputs are resident in HBM before timing starts (the structure is
uploaded once); each step is one C-ABI call: q upload, pair kernel, FP64 bin
reduction, contraction, I(q) download.

Prints one JSON line: metric pair.q evals/s (whole job), the dominant kernel's
roofline (timed with HIP events on its own stream inside the library), the FP64
GPU Debye baseline with the speedup and max relative error against it, and the
FP64 CPU oracle timed on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N_ATOMS, RADIUS, N_Q, SEED = 20000, 35.0, 500, 1
VALU_OPS_PER_PAIRQ = 4          # x = q r, sin x, * 1/(q r), accumulate
PEAK_VALU_OPS = 256 * 64 * 2.4e9  # CUs x lanes x clock (unpacked FP32 issue)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU baseline")
    ap.add_argument("--sin-mode", type=int, default=None)
    ap.add_argument("--q-tile", type=int, default=None)
    ap.add_argument("--j-chunk", type=int, default=None)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(nbytes: int) -> None:
    """One elementwise device copy of a known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE give the counter's bytes-per-KiB scale."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "read": nbytes, "written": nbytes}))


def cpu_baseline(structure, q, threads):
    """FP64 oracle (C restatement of scattering.rs:86-110) on a q subset."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_py as O
    atoms = O.Atoms(structure.xyz, structure.element, structure.group, structure.occupancy)
    sample = q[np.linspace(1, len(q) - 1, threads).astype(int)]
    O.debye_intensity(atoms, sample[:1], threads=1)  # warm-up of the loader
    t0 = time.perf_counter()
    O.debye_intensity(atoms, sample, threads=threads)
    dt = time.perf_counter() - t0
    n = len(structure)
    pairq = n * (n - 1) / 2 * len(sample)
    return {"value": pairq / dt, "unit": "pair.q evals/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle debye_intensity on the same 20,000-atom structure at "
                      f"{len(sample)} of the 500 q values ({pairq:.3g} pair.q, {dt:.1f} s); "
                      f"one q per thread"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not "
                                   "part of it (use --gpus 1)"}))
        sys.exit(2)
    import crabsaxs_amd as csx

    ctx = csx.default_context()
    tune = dict(q_tile=64, sin_mode=2, j_chunk=512)
    if args.sin_mode is not None:
        tune["sin_mode"] = args.sin_mode
    if args.q_tile is not None:
        tune["q_tile"] = args.q_tile
    if args.j_chunk is not None:
        tune["j_chunk"] = args.j_chunk
    ctx.set_tuning(**tune)

    s = csx.generate_globule(N_ATOMS, RADIUS, 1.0, seed=SEED)
    q = np.linspace(0.0, 1.0, N_Q)
    params = csx.SolventParams()
    s.handle(ctx)  # upload before timing
    F32, F64 = csx.Precision.FP32_COMPENSATED, csx.Precision.FP64_PLAIN

    for _ in range(args.warmup):
        out = csx.solvent_corrected_intensity(s, q, params, precision=F32, ctx=ctx)
    ctx.reset_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = csx.solvent_corrected_intensity(s, q, params, precision=F32, ctx=ctx)
    dt = time.perf_counter() - t0
    ctx.set_timing(False)
    launches, kernel_ms = ctx.kernel_time("debye_pairs_f32")
    pairq = N_ATOMS * (N_ATOMS - 1) / 2 * N_Q
    ms_per_step = dt / args.steps * 1e3
    value = pairq * args.steps / dt
    avg_kernel_s = kernel_ms / max(launches, 1) / 1e3
    achieved = pairq / avg_kernel_s * VALU_OPS_PER_PAIRQ

    result = {
        "metric": "pair.q evals/s (FP32-compensated Debye, <=1e-5 max rel err vs FP64 GPU Debye)",
        "value": value,
        "unit": "pair.q evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (compensated; fp64 cross-block reduction)",
        "data": "synthetic",
        "config": {"workload": "Debye precision study: seeded 20,000-atom globule, R=35 A, "
                               "q=linspace(0,1,500), default solvent, per-type S_ab tables",
                   "atoms": N_ATOMS, "q_points": N_Q, "pairs": N_ATOMS * (N_ATOMS - 1) // 2,
                   "tuning": tune},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_VALU_OPS / 1e12,
                     "unit": "T VALU-op/s", "frac": achieved / PEAK_VALU_OPS, "traffic": None,
                     "kernel": "debye_pairs_f32", "kernel_ms": avg_kernel_s * 1e3,
                     "ops_per_pairq": VALU_OPS_PER_PAIRQ},
    }
    if not args.no_fp64:
        ctx.reset_timing()
        ctx.set_timing(True)
        t1 = time.perf_counter()
        ref = csx.solvent_corrected_intensity(s, q, params, precision=F64, ctx=ctx)
        dt64 = time.perf_counter() - t1
        ctx.set_timing(False)
        _, ms64 = ctx.kernel_time("debye_plain_f64")
        err = np.abs(out - ref) / np.abs(ref)
        result["fp64_gpu"] = {"value": pairq / dt64, "unit": "pair.q evals/s",
                              "kernel_ms": ms64, "speedup": value / (pairq / dt64),
                              "max_rel_err": float(err.max()), "mean_rel_err": float(err.mean())}
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        result["cpu_baseline"] = cpu_baseline(s, q, threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
