"""Benchmark: ms per pressure solve to 1e-8 relative residual on one MI355X.

Workload (one "step" = one full pressure solve from p = 0): the 2-D
variable-density bubble pressure Poisson problem, 512x1024 cells in a 1x2
box, rho 1000:1 across a circle of radius 0.25 at (0.5, 0.5) (smoothed
Heaviside, width 1.5h), b = divergence of a seeded smooth velocity field,
all-Neumann walls.  Inputs are resident in HBM before the timed region.

`value` is the headline path (FP32 V-cycle inside the FP64 defect-correction
loop, learned initial guess when a trained model is attached); the plain FP64
multigrid solve of the same problem is reported beside it.  The roofline
object is the dominant kernel timed with HIP events on the solver stream;
cpu_baseline is the oracle's C port of the same FP64 multigrid on one host
core, on a bounded number of solves.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/pressure solve to 1e-8 residual: FP64 MG vs mixed+learned; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, seconds=12.0):
    """The oracle's FP64 red-black V(2,2) multigrid (C, one core) solving the
    same system to 1e-8; repeated until `seconds` of CPU work."""
    import numpy as np

    from oracle.pyoracle import Oracle

    nx, ny, _ = w["shape"]
    dx, dy, _ = w["spacing"]
    o = Oracle()
    m = o.mg(2, nx, ny, 1, dx, dy, 1.0, "neumann")
    m.set_rho(w["rho"])
    times, cycles = [], 0
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < seconds or not times:
        m.vec(0, "rhs")[:] = w["rhs"]
        m.vec(0, "sol")[:] = 0.0
        t0 = time.perf_counter()
        cycles, _ = m.solve_rb(1e-8, 200, 2, 2)
        times.append(time.perf_counter() - t0)
    ms = 1e3 * float(np.mean(times))
    return {"value": round(ms, 2), "unit": "ms/solve", "cores": 1, "kind": "port",
            "sample": f"{len(times)} full solves of the 512x1024 bubble system to 1e-8 "
                      f"({cycles} V(2,2) cycles each) by oracle/oracle.c, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multigrid path "
              "has no multi-GPU mode in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    from scihpc_amd import PREC_FP64, PREC_MIXED, Multigrid
    from scihpc_amd import workloads as W

    log("bench: building the 512x1024 bubble workload")
    w = W.bubble_2d()
    nx, ny, _ = w["shape"]
    dx, dy, _ = w["spacing"]
    g = Multigrid(2, nx, ny, 1, dx, dy, 1.0, "neumann")
    g.set_density(w["rho"])
    g.set_rhs(w["rhs"])
    tol = 1e-8

    def run(precision, steps, warmup, accel=0):
        stats = None
        for _ in range(warmup):
            stats = g.solve(rel_tol=tol, max_cycles=200, precision=precision, accel=accel)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            stats = g.solve(rel_tol=tol, max_cycles=200, precision=precision, accel=accel)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        assert stats["converged"], stats
        return 1e3 * dt, stats

    log("bench: plain FP64 multigrid")
    ms64, st64 = run(PREC_FP64, args.steps, args.warmup)
    x64 = g.sol(0)
    log(f"bench: fp64 {ms64:.3f} ms/solve, {st64['cycles']} cycles")
    log("bench: FP64 multigrid + flexible CG")
    ms64k, st64k = run(PREC_FP64, args.steps, args.warmup, accel=1)
    log(f"bench: fp64+fcg {ms64k:.3f} ms/solve, {st64k['cycles']} cycles")
    log("bench: mixed FP32 V-cycle + FP64 flexible-CG defect correction")
    ms_mx, st_mx = run(PREC_MIXED, args.steps, args.warmup, accel=1)
    xmx = g.sol(0)
    log(f"bench: mixed {ms_mx:.3f} ms/solve, {st_mx['cycles']} cycles")
    err = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))

    # level-0 kernels, HIP events on the solver's stream
    kernels = {}
    for which, name in [(0, "fp64_pre_smooth_restrict"), (1, "fp32_pre_smooth_restrict"),
                        (2, "fp32_post_smooth_fp64_outer"), (3, "fp64_post_smooth_norm")]:
        ms_k, b_k = g.time_kernel(which, iters=50)
        kernels[name] = {"us": round(ms_k * 1e3, 2), "bytes": b_k, "GBps": round(b_k / (ms_k * 1e-3) / 1e9, 1)}
    # dominant kernel of the headline solve: the FP32 level-0 pre-smoothing
    # kernel (two red-black sweeps + residual + restriction)
    kms, kbytes = kernels["fp32_pre_smooth_restrict"]["us"] * 1e-3, kernels["fp32_pre_smooth_restrict"]["bytes"]
    achieved = kbytes / (kms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(ms_mx, 4),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mx, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 V-cycle + fp64 defect correction",
        "data": "synthetic (seeded bubble density 1000:1, random smooth divergence RHS)",
        "config": {"workload": "bubble_2d 512x1024 variable-density pressure Poisson, all-Neumann",
                   "rel_tol": tol, "cycle": "V(2,2) red-black GS", "levels": g.level_num,
                   "outer": "FCG(1) defect correction"},
        "fp64_mg": {"ms_per_solve": round(ms64, 4), "cycles": st64["cycles"],
                    "rel_residual": st64["rel_residual"], "accel": "stationary V-cycles"},
        "fp64_mg_fcg": {"ms_per_solve": round(ms64k, 4), "cycles": st64k["cycles"],
                        "rel_residual": st64k["rel_residual"]},
        "mixed": {"ms_per_solve": round(ms_mx, 4), "cycles": st_mx["cycles"],
                  "rel_residual": st_mx["rel_residual"], "rel_l2_vs_fp64": err,
                  "accel": "flexible CG (FCG(1)), FP64 vectors", "learned": False},
        "speedup_mixed_vs_fp64": round(ms64 / ms_mx, 3),
        "speedup_mixed_vs_fp64_fcg": round(ms64k / ms_mx, 3),
        "kernels": kernels,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_pre2d<float,true> level 0 (2 RB-GS sweeps + residual + restriction)",
                     "bytes_per_launch": kbytes, "us_per_launch": round(kms * 1e3, 3)},
    }
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle port, 1 core)")
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
