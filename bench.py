"""Benchmark: ms per pressure solve to 1e-8 relative residual on one MI355X.

Workload (one "step" = one full pressure solve from p = 0): the 2-D
variable-density bubble pressure Poisson problem, 512x1024 cells in a 1x2
box, rho 1000:1 across a circle of radius 0.25 at (0.5, 0.5) (smoothed
Heaviside, width 1.5h), b = divergence of a seeded smooth velocity field,
all-Neumann walls.  Inputs are resident in HBM before the timed region; the
timed solves rotate over four such right-hand sides (four seeds), each
copied device to device into the solver at the start of its solve, as a
time loop hands every solve a new b.

`value` is the path the metric names: the FP32 V-cycle inside the FP64
flexible-CG defect correction, with the trained U-Net guess evaluated and
judged by the FP64 residual on every solve (kept only if it lowers it).
Beside it: plain FP64 multigrid (stationary V-cycles), FP64 + flexible CG,
and the mixed path without the network.  The roofline object is the
dominant kernel timed with HIP events on the solver stream; the U-Net's
MFMA rate is reported from its own timing hook; the 256^3 variable-density
3-D solves (the 1000:1 jump across a wavy plane and the constant-coefficient
Dirichlet problem) and the WENO5 + TVD-RK3 level-set step are reported as
well; and cpu_baseline is the oracle's C port of the same FP64 multigrid on
the host's threads and on one thread, beside the reference-faithful FPC
multigrid path, on a bounded number of solves.
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
MFMA_BF16_PEAK = 2500.0  # TFLOP/s, dense
# HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE) of `bench.py --kernel-only`, each scaled by what the
# counter reads for a 256 MiB device copy (FETCH x 2.0, WRITE x 1.0):
# profiles/pmc_traffic_k_pre2d_fp32_bubble_512x1024.md
PRE2D_FP32_TRAFFIC_BYTES = (4362.8 * 2.0 + 3072.0 * 1.0) * 1024
# the same for the FCG update fused into that kernel (the headline solve's
# level-0 pre-smoothing from the second iteration on):
# profiles/pmc_traffic_k_pre2d_fcg_bubble_512x1024.md
FCG_PRE2D_TRAFFIC_BYTES = (10757.0 * 2.0 + 13344.0 * 1.0) * 1024


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads for the multi-threaded CPU legs: OMP_NUM_THREADS when set
    (the GPU box sets it to its CPU share), else the cores this process may
    use, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(w, rhs_list, seconds=12.0):
    """CPU legs beside the GPU solve, on the host of the same run.

    value: the oracle's C port (oracle/oracle.c, g++-compatible gcc -O3
    -fopenmp) of the same FP64 red-black V(2,2) multigrid on the same A(rho)
    system, solving the same rotating right-hand sides to 1e-8, with the
    host's threads; beside it the same on one thread, and the
    reference-faithful path of SciHPC's fast_pressure_correction
    (steepest-descent relax, child-average restriction, slope
    prolongation, direct coarse solve; full cycles until the relative
    residual is 1e-8) on the constant-coefficient Neumann operator that
    path solves (it diverges on the 1000:1 A(rho) operator)."""
    import numpy as np

    from oracle.pyoracle import Oracle

    nx, ny, _ = w["shape"]
    dx, dy, _ = w["spacing"]
    o = Oracle()
    nthr = cpu_threads()

    def rb_leg(threads, budget):
        used = o.lib.orc_set_threads(threads)
        m = o.mg(2, nx, ny, 1, dx, dy, 1.0, "neumann")
        m.set_rho(w["rho"])
        times, cycles = [], 0
        t_start = time.perf_counter()
        k = 0
        while time.perf_counter() - t_start < budget or not times:
            m.vec(0, "rhs")[:] = rhs_list[k % len(rhs_list)]
            m.vec(0, "sol")[:] = 0.0
            t0 = time.perf_counter()
            cycles, _ = m.solve_rb(1e-8, 200, 2, 2)
            times.append(time.perf_counter() - t0)
            k += 1
        return used, 1e3 * float(np.mean(times)), len(times), cycles

    used, ms_n, cnt_n, cyc = rb_leg(nthr, 0.4 * seconds)
    _, ms_1, cnt_1, cyc1 = rb_leg(1, 0.4 * seconds)
    # reference-faithful FPC-mode solve (constant-coefficient Neumann)
    o.lib.orc_set_threads(nthr)
    m = o.mg(2, nx, ny, 1, dx, dy, 1.0, "neumann")
    m.vec(0, "rhs")[:] = rhs_list[0]
    m.vec(0, "sol")[:] = 0.0
    bn = float(np.linalg.norm(rhs_list[0]))
    t0 = time.perf_counter()
    ref_cycles = 0
    while o.lib.orc_mg_residual(m.h, 0) / bn > 1e-8 and ref_cycles < 50:
        o.lib.orc_mg_fullcycle_ref(m.h)
        ref_cycles += 1
    ms_ref = (time.perf_counter() - t0) * 1e3
    ref_rr = o.lib.orc_mg_residual(m.h, 0) / bn
    o.lib.orc_set_threads(1)
    return {"value": round(ms_n, 2), "unit": "ms/solve", "cores": used, "kind": "port",
            "cpu_model": cpu_model(), "threads": used,
            "sample": f"{cnt_n} full solves of the 512x1024 bubble A(rho) system to 1e-8 ({cyc} V(2,2) cycles "
                      f"each), rotating over {len(rhs_list)} right-hand sides, by oracle/oracle.c "
                      f"(gcc -O3 -fopenmp, {used} threads)",
            "one_thread": {"ms_per_solve": round(ms_1, 2), "solves": cnt_1, "cycles": cyc1},
            "reference_faithful_fpc": {
                "ms_per_solve": round(ms_ref, 1), "full_cycles": ref_cycles, "rel_residual": ref_rr,
                "threads": used,
                "path": "steepest-descent V(10,10) full cycles (multigrid.cpp:57-79, multigrid_base.cpp:16-234) "
                        "on the constant-coefficient Neumann operator of the FPC solve, same b, to 1e-8"}}


def bench_weno(reps=10):
    """WENO5-Z level-set TVD-RK3 step (weno_solver.cpp, runge_kutta.cpp:13-49,
    source.cpp:8-40 with the WENO call) at the 1024x512 dam-break shape:
    device time per step on a resident level set against the minimal
    Shu-Osher traffic (7 + 3d) * 8 B per cell."""
    from scihpc_amd import LevelSet

    wl = momentum_workload()
    nx, ny = wl["shape"]
    h = wl["h"]
    ls = LevelSet(2, nx, ny, 1, h, h, 1.0, ["neumann"] * 6)
    ls.set_phi(wl["phi"])
    ls.set_velocity(wl["u"], wl["v"])
    ms, byts = ls.time_step(0.25 * h, iters=reps)
    ls.close()
    gbs = byts / (ms * 1e-3) / 1e9
    return {"workload": f"dam break {nx}x{ny} level set, WENO5-Z + TVD-RK3", "ms_per_step": round(ms, 4),
            "bytes_per_step": byts, "GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def cpu_weno_baseline():
    """The oracle's WENO5 TVD-RK3 step (one core) at the same shape."""
    import numpy as np

    from oracle.pyoracle import Oracle

    wl = momentum_workload()
    nx, ny = wl["shape"]
    h = wl["h"]
    o = Oracle()
    o.lib.orc_set_threads(1)
    phi = wl["phi"].copy()
    t0 = time.perf_counter()
    o.tvd_rk3(2, (nx, ny, 1), (h, h, 1.0), 0.25 * h, ["neumann"] * 4, [0.0] * 4, phi, wl["u"], wl["v"],
              np.zeros_like(phi), 1)
    return {"ms_per_step": round((time.perf_counter() - t0) * 1e3, 2), "cores": 1, "kind": "port",
            "sample": f"one step at {nx}x{ny}"}


def kernel_only(args):
    """21 launches each of the FP32 level-0 pre-smoothing kernel of the
    bubble solve and of its form with the FCG update fused in (after one
    mixed FCG solve), then 5 copies of 256 MiB with torch: reading
    and writing a known byte count calibrates the FETCH_SIZE / WRITE_SIZE
    counters."""
    import torch

    from scihpc_amd import PREC_MIXED, Multigrid
    from scihpc_amd import workloads as W

    w = W.bubble_2d()
    nx, ny, _ = w["shape"]
    dx, dy, _ = w["spacing"]
    g = Multigrid(2, nx, ny, 1, dx, dy, 1.0, "neumann")
    g.set_density(w["rho"])
    g.set_rhs(w["rhs"])
    g.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1)
    ms, b = g.time_kernel(1, iters=20)
    ms6, b6 = g.time_kernel(6, iters=20)
    a = torch.empty(64 * 2 ** 20, dtype=torch.float32, device="cuda")
    c = torch.ones_like(a)
    for _ in range(5):
        a.copy_(c)
    torch.cuda.synchronize()
    print(json.dumps({"kernel": "k_pre2d<float,true,32,16>", "us": ms * 1e3, "bytes": b,
                      "fused_kernel": "k_pre2d_fcg<true,32,16>", "fused_us": ms6 * 1e3, "fused_bytes": b6,
                      "calibration_copy_bytes_each_way": a.numel() * 4}), flush=True)


def bench_3d(args):
    """256^3 variable-density solve (1000:1 across z = 0.5 + 0.05 sin sin),
    mixed vs FP64, and the level-0 kernels with HIP events."""
    import numpy as np
    import torch

    from scihpc_amd import PREC_FP64, PREC_MIXED, Multigrid
    from scihpc_amd import workloads as W

    n = 256
    log("bench: building the 256^3 plane workload")
    w = W.poisson_3d(n, variable=True)
    g = Multigrid(3, n, n, n, 1 / n, 1 / n, 1 / n, "neumann")
    g.set_density(w["rho"])
    g.set_rhs(w["rhs"])
    res = {"workload": "plane_3d 256^3 variable-density pressure Poisson, all-Neumann", "rel_tol": 1e-8}
    xs = {}
    for name, prec, acc, lag in [("fp64_mg", PREC_FP64, 0, 0), ("fp64_mg_fcg", PREC_FP64, 1, 0),
                                 ("mixed", PREC_MIXED, 1, 0), ("fp64_mg_lagged", PREC_FP64, 0, 1),
                                 ("mixed_lagged", PREC_MIXED, 1, 1)]:
        g.solve(rel_tol=1e-8, max_cycles=100, precision=prec, accel=acc, lagged_check=lag)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            st = g.solve(rel_tol=1e-8, max_cycles=100, precision=prec, accel=acc, lagged_check=lag)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        assert st["converged"], st
        xs[name] = g.sol(0)
        res[name] = {"ms_per_solve": round(ms, 3), "cycles": st["cycles"], "rel_residual": st["rel_residual"]}
        log(f"bench: 3-D {name} {ms:.2f} ms/solve, {st['cycles']} cycles")
    res["mixed"]["rel_l2_vs_fp64"] = float(np.linalg.norm(xs["mixed"] - xs["fp64_mg"]) /
                                           np.linalg.norm(xs["fp64_mg"]))
    res["speedup_mixed_vs_fp64"] = round(res["fp64_mg"]["ms_per_solve"] / res["mixed"]["ms_per_solve"], 3)
    res["speedup_mixed_lagged_vs_fp64"] = round(res["fp64_mg"]["ms_per_solve"] /
                                                res["mixed_lagged"]["ms_per_solve"], 3)
    kern = {}
    for which, name in [(0, "fp64_pre_smooth_restrict"), (1, "fp32_pre_smooth_restrict"),
                        (2, "fp32_post_smooth_fp64_outer"), (3, "fp64_post_smooth_norm"),
                        (4, "fp64_prolong_add"), (5, "fp64_residual_pass")]:
        ms_k, b_k = g.time_kernel(which, iters=5 if which < 4 else 20)
        kern[name] = {"us": round(ms_k * 1e3, 1), "bytes": b_k, "GBps": round(b_k / (ms_k * 1e-3) / 1e9, 1),
                      "frac": round(b_k / (ms_k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    res["kernels"] = kern
    res["kernels_note"] = ("each mode relaunches one kernel over the same arrays, so part of them stays in the "
                           "256 MB last-level cache between launches (fp64_prolong_add exceeds the device-copy "
                           "ceiling this way); inside a solve the arrays are evicted between uses")
    g.close()
    del g
    # the constant-coefficient Dirichlet half of the configuration: f from
    # seeded random Fourier modes (k <= 8, N(0,1)/|k|^2)
    log("bench: building the 256^3 Dirichlet workload")
    wd = W.poisson_3d(n, variable=False)
    gd = Multigrid(3, n, n, n, 1 / n, 1 / n, 1 / n, "dirichlet")
    gd.set_rhs(wd["rhs"])
    dres = {"workload": "poisson_3d 256^3 constant-coefficient Dirichlet, random Fourier f", "rel_tol": 1e-8}
    xd = {}
    for name, prec, acc in [("fp64_mg", PREC_FP64, 0), ("fp64_mg_fcg", PREC_FP64, 1), ("mixed", PREC_MIXED, 1)]:
        gd.solve(rel_tol=1e-8, max_cycles=100, precision=prec, accel=acc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            st = gd.solve(rel_tol=1e-8, max_cycles=100, precision=prec, accel=acc)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        assert st["converged"], st
        xd[name] = gd.sol(0)
        dres[name] = {"ms_per_solve": round(ms, 3), "cycles": st["cycles"], "rel_residual": st["rel_residual"]}
        log(f"bench: 3-D Dirichlet {name} {ms:.2f} ms/solve, {st['cycles']} cycles")
    gd.solve(rel_tol=1e-12, max_cycles=100, precision=PREC_FP64, accel=1)
    x_ref = gd.sol(0)
    for name in xd:
        dres[name]["rel_l2_vs_fp64_1e-12"] = float(np.linalg.norm(xd[name] - x_ref) / np.linalg.norm(x_ref))
    dres["speedup_mixed_vs_fp64"] = round(dres["fp64_mg"]["ms_per_solve"] / dres["mixed"]["ms_per_solve"], 3)
    gd.close()
    res["dirichlet_fourier"] = dres
    return res


def bench_learned3d(reps=3, problems=3):
    """Config 5's 3-D half: held-out 128^3 bubbles (sphere of radius 0.2 at a
    seeded position, 1000:1) solved by mixed FCG to 1e-8 with and without
    the 3-D U-Net guess (network on level 2, 32^3), judged on every solve
    and forced; the guess's energy-norm error against the FP64 solution."""
    import numpy as np
    import torch

    from scihpc_amd import PREC_FP64, PREC_MIXED, Multigrid
    from scihpc_amd import learned as LRN

    path = os.path.join(LRN.WEIGHTS_DIR, "unet_bubble3d_128_level2.npz")
    if not os.path.exists(path):
        return {"attached": False, "note": "no 3-D weights (python -m scihpc_amd.learned --data bubble3d)"}
    net, meta = LRN.load_net(path)
    n = 128
    g = Multigrid(3, n, n, n, 1 / n, 1 / n, 1 / n, "neumann")
    g.attach_learned(net, meta["level"])
    rows = []
    for cfg in LRN.bubble3d_configs(problems, LRN.EVAL_SEED):
        w = LRN.make_problem(cfg, n)
        g.set_density(w["rho"])
        g.set_rhs(w["rhs"])
        row = {"center": [round(v, 3) for v in cfg["center"]]}
        for name, learned in (("plain", 0), ("learned_judged", 3), ("learned_forced", 2)):
            g.solve(rel_tol=1e-8, max_cycles=100, precision=PREC_MIXED, accel=1, use_learned=learned)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                st = g.solve(rel_tol=1e-8, max_cycles=100, precision=PREC_MIXED, accel=1, use_learned=learned)
            torch.cuda.synchronize()
            assert st["converged"], st
            row[name] = {"ms_per_solve": round(1e3 * (time.perf_counter() - t0) / reps, 3), "cycles": st["cycles"]}
            if learned == 3:
                row[name]["accepted"] = st["learned_accepted"]
                row[name]["guess_rel_residual"] = st["learned_residual"]
        q = LRN.guess_quality(g, [w])[0]
        row["guess_energy_error"] = round(q["energy"], 4)
        rows.append(row)
        log(f"bench: 3-D learned {row}")
    g.close()
    net.close()
    return {"attached": True, "weights": os.path.relpath(path, ROOT), "net_level": meta["level"],
            "net_grid": [meta["width"], meta["height"], meta["depth"]], "training": meta["meta"],
            "workload": "bubble_3d 128^3, radius 0.2, rho 1000:1, held-out positions", "rel_tol": 1e-8,
            "problems": rows}


def bench_learned_dambreak(reps=3):
    """Config 5's 2-D half: the dam-break network (trained on the solver's
    own 1024x512 dam-break snapshots, steps 10-1040) on two held-out steps
    (655: an ordinary step; 1310: a hard step before the flow breaks down):
    mixed FCG to 1e-8 from p = 0, from the network's guess, and with the
    guess judged inside the solve."""
    import torch

    from scihpc_amd import PREC_MIXED, Multigrid
    from scihpc_amd import learned as LRN

    path = os.path.join(LRN.WEIGHTS_DIR, "unet_dambreak_1024x512_level3.npz")
    if not (os.path.exists(path) and os.path.exists(LRN.DAMBREAK_FIXTURE)):
        return {"attached": False}
    net, meta = LRN.load_net(path)
    g = Multigrid(2, 1024, 512, 1, 4 / 1024, 2 / 512, 1.0, "neumann")
    g.attach_learned(net, meta["level"])
    rows = {}
    for step, w in sorted(LRN.dambreak_fixture_problems().items()):
        g.set_density(w["rho"])
        g.set_rhs(w["rhs"])
        row = {}
        for name, learned in (("plain", 0), ("from_guess", 2), ("judged", 3)):
            g.solve(rel_tol=1e-8, max_cycles=200, precision=PREC_MIXED, accel=1, use_learned=learned)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                st = g.solve(rel_tol=1e-8, max_cycles=200, precision=PREC_MIXED, accel=1, use_learned=learned)
            torch.cuda.synchronize()
            assert st["converged"], st
            row[name] = {"ms_per_solve": round(1e3 * (time.perf_counter() - t0) / reps, 3), "cycles": st["cycles"]}
        rows[str(step)] = row
        log(f"bench: dam-break learned step {step}: {row}")
    g.close()
    net.close()
    return {"attached": True, "weights": os.path.relpath(path, ROOT), "net_level": meta["level"],
            "training": meta["meta"], "workload": "dam break 1024x512 A(rho) pressure systems, held-out steps",
            "rel_tol": 1e-8, "steps": rows}


def bench_unet(iters=10):
    """The learned preconditioner's network (4 levels, 16/32/64/128
    channels, seed 3) at 256x512, batch 4, on random samples: ms and
    TFLOP/s of a forward pass and of a training step (forward, backward,
    Adam) on the matrix cores."""
    import numpy as np

    from scihpc_amd import UNet

    nt = UNet(512, 256, 4, seed=3)
    rng = np.random.default_rng(5)
    nt.out_scale = 1.0
    nt.set_samples(rng.standard_normal((4, 512 * 256, 2)).astype(np.float32),
                   rng.standard_normal((4, 512 * 256)).astype(np.float32))
    ms_tr, fl_tr = nt.time(4, train=True, iters=iters)
    ms_fw, fl_fw = nt.time(4, train=False, iters=iters)
    nt.close()
    return {"grid": [256, 512], "batch": 4,
            "train_step_ms": round(ms_tr, 3), "train_tflops": round(fl_tr / (ms_tr * 1e-3) / 1e12, 2),
            "forward_ms": round(ms_fw, 3), "forward_tflops": round(fl_fw / (ms_fw * 1e-3) / 1e12, 2),
            "mfma_peak_tflops": MFMA_BF16_PEAK,
            "mfma_frac_train": round(fl_tr / (ms_tr * 1e-3) / 1e12 / MFMA_BF16_PEAK, 4),
            "mfma_frac_forward": round(fl_fw / (ms_fw * 1e-3) / 1e12 / MFMA_BF16_PEAK, 4)}


def bench_snapshot(torch, nx=512, ny=1024, reps=5):
    """Legacy-VTK snapshot of the 512x1024 grid as the reference's drivers
    write it (three scalars and one vector per file, vtkWriter.cpp:18-111):
    the reorder + byte-swap kernel alone against HBM and the whole snapshot
    from device-resident fields to a closed file.  The fields and the file's
    bytes ride along (underscore keys) for the CPU baseline leg."""
    import tempfile

    import numpy as np

    from scihpc_amd import VtkWriter, padded_shape

    shp = padded_shape(2, nx, ny)
    cells = nx * ny
    h = (1.0 / nx, 2.0 / ny)
    dev = [torch.randn(int(np.prod(shp)), dtype=torch.float64, device="cuda") for _ in range(6)]
    torch.cuda.synchronize()
    out = {"grid": [nx, ny], "fields": "phi, density, pressure (scalars) + nvel (vector)"}
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        wr = VtkWriter(2, (nx, ny), (0.0, 0.0), h, os.path.join(d, "snap"))
        ms1 = wr.time_pack(1, iters=50)
        ms3 = wr.time_pack(3, iters=50)
        out["pack_scalar_us"] = round(ms1 * 1e3, 2)
        out["pack_scalar_GBps"] = round(16.0 * cells / (ms1 * 1e-3) / 1e9, 1)
        out["pack_vector_us"] = round(ms3 * 1e3, 2)
        out["pack_vector_GBps"] = round(48.0 * cells / (ms3 * 1e-3) / 1e9, 1)
        file_bytes = 0
        t0 = time.perf_counter()
        for r in range(reps):
            path = wr.begin(r)
            for name, t in zip(("phi", "density", "pressure"), dev[:3]):
                wr.add_scalar_device(name, t.data_ptr())
            wr.add_vector_device("nvel", *(t.data_ptr() for t in dev[3:]))
            wr.close()
            file_bytes = os.path.getsize(path)
        ms_snap = (time.perf_counter() - t0) / reps * 1e3
        out["snapshot_ms"] = round(ms_snap, 3)
        out["file_MB"] = round(file_bytes / 1e6, 3)
        out["snapshot_MBps"] = round(file_bytes / (ms_snap * 1e-3) / 1e6, 1)
        out["_fields"] = [t.cpu().numpy() for t in dev]
        out["_file"] = open(path, "rb").read()
        wr.close_handle()
    return out


MOM_SHAPE = (1024, 512)
N_RHS = 4  # right-hand sides the timed 2-D solves rotate over


def momentum_workload(nx=MOM_SHAPE[0], ny=MOM_SHAPE[1]):
    """2-D dam break on [0,4] x [0,2] (1024 x 512 cells): phi = signed distance
    to the water column [0,1] x [0,2], air/water ratios, gravity, no slip,
    a seeded smooth velocity field with its wall values."""
    import numpy as np

    from scihpc_amd import padded_shape

    h = 4.0 / nx
    shp = padded_shape(2, nx, ny)
    I, J = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), indexing="ij")
    x, y = (I - 3 + 0.5) * h, (J - 3 + 0.5) * h
    phi = (1.0 - x).ravel()  # column of full height: distance to x = 1
    rng = np.random.default_rng(41)
    a = rng.standard_normal(4)
    u = (0.1 * (a[0] * np.sin(np.pi * x / 4) * np.cos(np.pi * y / 2) + a[1])).ravel()
    v = (0.1 * (a[2] * np.cos(np.pi * x / 4) * np.sin(np.pi * y / 2) + a[3])).ravel()
    L = 0.0571
    Uc = np.sqrt(9.81 * L)
    prm = dict(dt=0.01 * h, reynolds=1000.0 * Uc * L / 1.137e-3, froude=1.0, weber=-1.0, ls_width=1.5 * h,
               density_ratio=1.226e-3, viscosity_ratio=1.78e-5 / 1.137e-3)
    return dict(shape=(nx, ny), h=h, phi=phi, u=u, v=v, w=np.zeros_like(u), prm=prm)


# algorithmic bytes per padded cell of one 2-D predictor step without surface
# tension: level-set properties 6 arrays, per direction the face transfer 4
# and the source 9, AB2 12 (8-byte values, each array once per pass)
MOM_BYTES_PER_CELL_2D = 8 * (6 + 2 * (4 + 9) + 12)


def bench_hbm_ceiling(torch, mib=256, reps=20):
    """Measured streaming ceilings of this GPU, for reading the roofline
    fractions: a device-to-device copy of `mib` MiB (read + write) and a
    three-read one-write FP64 elementwise pass (torch.addcmul), HIP events."""
    n = mib * 2 ** 20 // 8
    a, b, c = (torch.randn(n, dtype=torch.float64, device="cuda") for _ in range(3))
    o = torch.empty_like(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    for name, fn, arrays in (("copy", lambda: o.copy_(a), 2), ("three_read_one_write", lambda: torch.addcmul(
            a, b, c, out=o), 4)):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        s = e0.elapsed_time(e1) * 1e-3 / reps
        out[name + "_GBps"] = round(arrays * 8 * n / s / 1e9, 1)
    out["bytes_per_array"] = 8 * n
    del a, b, c, o
    return out


def bench_momentum(iters=20):
    """find_source_sec + find_intermediate_velocity at the dam-break shape:
    the device time of one step on resident fields against HBM."""
    import numpy as np

    from scihpc_amd import Projection

    wl = momentum_workload()
    nx, ny = wl["shape"]
    pj = Projection(2, nx, ny, 1, wl["h"], wl["h"], 1.0, ["no_slip"] * 6)
    prm = Projection.flow_params(**wl["prm"])
    z = np.zeros_like(wl["u"])
    pj.momentum_step(prm, wl["phi"], wl["u"], wl["v"], wl["w"], z, z, z)
    ms = pj.time_momentum(prm, iters=iters)
    cells = int(np.prod(pj.shape))
    byts = MOM_BYTES_PER_CELL_2D * cells
    out = {"workload": f"dam break {nx}x{ny}, air/water, gravity, no slip", "ms_per_step": round(ms, 4),
           "padded_cells": cells, "bytes_per_step": byts, "GBps": round(byts / (ms * 1e-3) / 1e9, 1),
           "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    pj.close()
    return out


def cpu_momentum_baseline():
    """The oracle's momentum step (one core) at the same shape."""
    import numpy as np

    from oracle.pyoracle import Oracle

    wl = momentum_workload()
    nx, ny = wl["shape"]
    p = wl["prm"]
    prm8 = np.array([p["dt"], p["reynolds"], p["froude"], p["weber"], p["ls_width"], p["density_ratio"],
                     p["viscosity_ratio"], 1.0])
    arrs = [wl["u"].copy(), wl["v"].copy(), wl["w"].copy()] + [np.zeros_like(wl["u"]) for _ in range(7)]
    o = Oracle()
    t0 = time.perf_counter()
    o.lib.orc_momentum_step(2, nx, ny, 1, wl["h"], wl["h"], 1.0, prm8, np.full(6, 3, np.int32), wl["phi"], *arrs)
    return {"ms_per_step": round((time.perf_counter() - t0) * 1e3, 2), "cores": 1, "kind": "port",
            "sample": f"one predictor step at {nx}x{ny}"}


def bench_interface(reps=5):
    """Interface maintenance at the dam-break shape (1024 x 512): one UCCD
    TVD-RK3 step, the 5 mpls Euler steps of a time step and one
    redistancing RK3 step, device-resident phi; with the sequential line
    solver (the default, bit-identical to the reference) and with the
    partitioned one."""
    from scihpc_amd import LevelSet

    wl = momentum_workload()
    nx, ny = wl["shape"]
    h = wl["h"]
    out = {"workload": f"dam break {nx}x{ny} level set"}
    for solver in ("sequential", "partitioned"):
        ls = LevelSet(2, nx, ny, 1, h, h, 1.0, ["neumann"] * 6)
        ls.set_line_solver(solver)
        ls.set_phi(wl["phi"])
        ls.set_velocity(wl["u"], wl["v"])
        ls.set_scheme("uccd")
        dt, lw, ratio = 0.01 * h, 1.5 * h, 1.226e-3
        ms_rk3, _ = ls.time_step(dt, iters=reps)
        m0 = ls.mass(lw, ratio)
        ls.mpls(dt, lw, ratio, m0, count=1)
        t0 = time.perf_counter()
        for _ in range(reps):
            ls.mpls(dt, lw, ratio, m0, count=5)
        ms_mpls = (time.perf_counter() - t0) / reps * 1e3
        ls.find_sign(lw)
        ls.find_delta(lw)
        ls.redistance(dt, 0.25 * h, 1)
        t0 = time.perf_counter()
        ls.redistance(dt, 0.25 * h, reps)
        ms_red = (time.perf_counter() - t0) / reps * 1e3
        ls.close()
        row = {"uccd_rk3_step_ms": round(ms_rk3, 4), "mpls_5_steps_ms": round(ms_mpls, 4),
               "redistance_rk3_step_ms": round(ms_red, 4)}
        if solver == "sequential":
            out.update(row)
        else:
            out["partitioned_line_solver"] = row
    return out


def cpu_interface_baseline():
    """The oracle's UCCD RK3 step, one mpls step and one redistancing step
    (one core) at the same shape."""
    import numpy as np

    from oracle.pyoracle import Oracle

    wl = momentum_workload()
    nx, ny = wl["shape"]
    h = wl["h"]
    o = Oracle()
    T = wl["phi"].size
    types, vals = np.zeros(6, np.int32), np.zeros(6)
    dt, lw, ratio = 0.01 * h, 1.5 * h, 1.226e-3
    phi = wl["phi"].copy()
    t0 = time.perf_counter()
    o.lib.orc_tvd_rk3_uccd(2, nx, ny, 1, h, h, 1.0, dt, types, vals, phi, wl["u"], wl["v"], wl["w"], 1)
    t_rk3 = time.perf_counter() - t0
    m0 = o.lib.orc_lsf_mass(2, nx, ny, 1, h * h, lw, ratio, 1, phi)
    delta = np.zeros(T)
    t0 = time.perf_counter()
    o.lib.orc_mpls_euler(2, nx, ny, 1, h, h, 1.0, dt, lw, ratio, 1, m0, types, vals, phi, 1, delta.ctypes.data)
    t_mpls = time.perf_counter() - t0
    sign = np.zeros(T)
    o.lib.orc_find_sign(T, phi, lw, sign)
    t0 = time.perf_counter()
    o.lib.orc_redistance_rk3(2, nx, ny, 1, h, h, 1.0, dt, 0.25 * h, sign, delta, types, vals, phi, 1)
    t_red = time.perf_counter() - t0
    return {"uccd_rk3_step_ms": round(t_rk3 * 1e3, 2), "mpls_5_steps_ms": round(5 * t_mpls * 1e3, 2),
            "redistance_rk3_step_ms": round(t_red * 1e3, 2), "cores": 1, "kind": "port",
            "sample": f"one step each at {nx}x{ny} (mpls: one step, x5)"}


def cpu_snapshot_baseline(snap):
    """The oracle's C writer (one core) on the same fields: ms per file and
    whether its bytes equal the GPU writer's file."""
    import tempfile

    from oracle.pyoracle import Oracle

    nx, ny = snap["grid"]
    f = snap["_fields"]
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        path = os.path.join(d, "cpu.vtk")
        t0 = time.perf_counter()
        Oracle().vtk_write(path, 2, (nx, ny), (0.0, 0.0, 0.0), (1.0 / nx, 2.0 / ny, 0.0),
                           list(zip(("phi", "density", "pressure"), f[:3])), [("nvel", tuple(f[3:]))])
        ms = (time.perf_counter() - t0) * 1e3
        same = open(path, "rb").read() == snap["_file"]
    return {"ms_per_file": round(ms, 3), "cores": 1, "identical_to_gpu_file": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-3d", action="store_true", help="skip the 256^3 3-D solves")
    ap.add_argument("--only-3d", action="store_true", help="run only the 256^3 3-D solves (profiling)")
    ap.add_argument("--no-unet", action="store_true", help="skip the U-Net training-step timing")
    ap.add_argument("--no-snapshot", action="store_true", help="skip the VTK snapshot timing")
    ap.add_argument("--no-momentum", action="store_true", help="skip the momentum predictor timing")
    ap.add_argument("--no-interface", action="store_true", help="skip the interface-maintenance timing")
    ap.add_argument("--no-weno", action="store_true", help="skip the WENO5 TVD-RK3 timing")
    ap.add_argument("--no-learned3d", action="store_true", help="skip the 128^3 learned-guess row")
    ap.add_argument("--only-learned3d", action="store_true", help="run only the 128^3 learned-guess row")
    ap.add_argument("--only-unet", action="store_true", help="run only the U-Net timing row (profiling)")
    ap.add_argument("--only-rows", action="store_true",
                    help="run only the rows around the solve (momentum predictor, interface maintenance, "
                         "VTK snapshot) and print their JSON (profiling)")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only launch the dominant kernel (and a 256 MiB copy for calibration): "
                         "the command profiled with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multigrid path "
              "has no multi-GPU mode in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    from scihpc_amd import PREC_FP64, PREC_MIXED, Multigrid, UNet
    from scihpc_amd import learned as LRN
    from scihpc_amd import workloads as W

    if args.kernel_only:
        kernel_only(args)
        return
    if args.only_3d:
        out3 = bench_3d(args)
        # the known-size device copy that calibrates PMC byte counters
        out3["measured_ceiling"] = bench_hbm_ceiling(torch)
        print(json.dumps(out3), flush=True)
        return
    if args.only_learned3d:
        print(json.dumps(bench_learned3d()), flush=True)
        return
    if args.only_unet:
        print(json.dumps(bench_unet()), flush=True)
        return
    if args.only_rows:
        rows = {"momentum": bench_momentum(), "interface": bench_interface(), "weno_rk3": bench_weno(),
                "snapshot": bench_snapshot(torch)}
        rows["snapshot"].pop("_fields")
        rows["snapshot"].pop("_file")
        print(json.dumps(rows), flush=True)
        return
    log("bench: building the 512x1024 bubble workload")
    w = W.bubble_2d()
    nx, ny, _ = w["shape"]
    dx, dy, _ = w["spacing"]
    g = Multigrid(2, nx, ny, 1, dx, dy, 1.0, "neumann")
    g.set_density(w["rho"])
    g.set_rhs(w["rhs"])
    tol = 1e-8
    # the timed solves rotate over right-hand sides of the same bubble
    # (divergences of four seeded velocity fields), resident on the device:
    # each solve starts by copying its b into the solver (device to
    # device), as a time loop hands a new b to every solve, so lagged
    # convergence checks pay their adaptation to changing iteration counts
    rhs_host = [w["rhs"]] + [W.bubble_2d(seed=W.SEED + 2 + 100 * k)["rhs"] for k in range(1, N_RHS)]
    rhs_dev = [torch.from_numpy(r).to("cuda") for r in rhs_host]
    torch.cuda.synchronize()
    turn = [0]

    def next_rhs():
        g.set_rhs_device(rhs_dev[turn[0] % N_RHS].data_ptr())
        turn[0] += 1

    def run(precision, steps, warmup, accel=0, learned=False, lagged=0):
        stats = None
        turn[0] = 0
        for _ in range(warmup):
            next_rhs()
            stats = g.solve(rel_tol=tol, max_cycles=200, precision=precision, accel=accel, use_learned=learned,
                            lagged_check=lagged)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cyc = []
        for _ in range(steps):
            next_rhs()
            stats = g.solve(rel_tol=tol, max_cycles=200, precision=precision, accel=accel, use_learned=learned,
                            lagged_check=lagged)
            assert stats["converged"], stats
            cyc.append(stats["cycles"])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        stats["cycles_per_solve"] = cyc
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
    # the same solves with lagged convergence checks: the host reads the
    # residual only from one iteration before the last solve's count
    log("bench: lagged convergence checks (FP64 stationary, mixed FCG)")
    ms64L, st64L = run(PREC_FP64, args.steps, args.warmup, lagged=1)
    ms_mxL, st_mxL = run(PREC_MIXED, args.steps, args.warmup, accel=1, lagged=1)
    errL = float(np.linalg.norm(g.sol(0) - x64) / np.linalg.norm(x64))
    log(f"bench: lagged fp64 {ms64L:.3f} ms/solve ({st64L['cycles']} cycles), mixed {ms_mxL:.3f} ms/solve "
        f"({st_mxL['cycles']} cycles)")

    learned = {"attached": False}
    ms_ml, st_ml = ms_mxL, st_mxL
    ms_ml_each = ms_mx
    if os.path.exists(LRN.DEFAULT_WEIGHTS):
        net, meta = LRN.load_net()
        g.attach_learned(net, meta["level"])
        log("bench: mixed + learned guess judged on every solve")
        ms_every, st_every = run(PREC_MIXED, args.steps, args.warmup, accel=1, learned=3)
        def backoff(lagged):
            g.attach_learned(net, meta["level"])  # fresh back-off state
            turn[0] = 0
            for _ in range(args.warmup):
                next_rhs()
                g.solve(rel_tol=tol, max_cycles=200, precision=PREC_MIXED, accel=1, use_learned=1,
                        lagged_check=lagged)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            skipped = accepted = 0
            for _ in range(args.steps):
                next_rhs()
                st = g.solve(rel_tol=tol, max_cycles=200, precision=PREC_MIXED, accel=1, use_learned=1,
                             lagged_check=lagged)
                skipped += st["learned_skipped"]
                accepted += st["learned_accepted"]
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / args.steps * 1e3
            assert st["converged"], st
            return ms, st, skipped, accepted

        log("bench: mixed + learned guess (FP64-judged, back-off after rejections)")
        ms_ml_each, _, _, _ = backoff(0)
        log("bench: the same with lagged convergence checks")
        ms_ml, st_ml, skipped, accepted = backoff(1)
        xml = g.sol(0)
        _, st_forced = run(PREC_MIXED, 1, 0, accel=1, learned=2)
        log("bench: mixed + learned coarse-level correction of the defect (FP64-judged)")
        turn[0] = 0
        for _ in range(args.warmup):
            next_rhs()
            g.solve(rel_tol=tol, max_cycles=200, precision=PREC_MIXED, accel=1, learned_correction=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        corr_kept = 0
        for _ in range(args.steps):
            next_rhs()
            st_corr = g.solve(rel_tol=tol, max_cycles=200, precision=PREC_MIXED, accel=1, learned_correction=1)
            corr_kept += st_corr["correction_accepted"]
        torch.cuda.synchronize()
        ms_corr = (time.perf_counter() - t0) / args.steps * 1e3
        assert st_corr["converged"], st_corr
        st_corr_forced = g.solve(rel_tol=tol, max_cycles=200, precision=PREC_MIXED, accel=1, learned_correction=2)
        f, t, _ = g.learned_features(meta["level"], with_target=True)
        net.set_samples(f[None], t[None])
        ms_inf, fl_inf = net.time(1, train=False, iters=20)
        learned = {"attached": True, "weights": os.path.relpath(LRN.DEFAULT_WEIGHTS, ROOT),
                   "net_level": meta["level"], "net_grid": [meta["width"], meta["height"]],
                   "training": meta["meta"], "ms_per_solve": round(ms_ml, 4), "cycles": st_ml["cycles"],
                   "lagged_check": True, "ms_per_solve_checked_every_iteration": round(ms_ml_each, 4),
                   "timed_solves": args.steps, "network_run_on": args.steps - skipped,
                   "guesses_accepted": accepted,
                   "every_solve": {"ms_per_solve": round(ms_every, 4), "cycles": st_every["cycles"],
                                   "accepted": st_every["learned_accepted"],
                                   "rejected": st_every["learned_rejected"],
                                   "guess_rel_residual": st_every["learned_residual"]},
                   "cycles_if_guess_forced": st_forced["cycles"],
                   "correction": {"ms_per_solve": round(ms_corr, 4), "cycles": st_corr["cycles"],
                                  "kept": corr_kept, "of": args.steps,
                                  "defect_rel_residual": st_corr["correction_residual"],
                                  "cycles_if_forced": st_corr_forced["cycles"]},
                   "rel_l2_vs_fp64": float(np.linalg.norm(xml - x64) / np.linalg.norm(x64)),
                   "inference_us": round(ms_inf * 1e3, 2),
                   "inference_mfma_tflops": round(fl_inf / (ms_inf * 1e-3) / 1e12, 3)}
        log(f"bench: mixed+learned {ms_ml:.3f} ms/solve, lagged checks ({ms_ml_each:.3f} checking every "
            f"iteration; network on {args.steps - skipped} of {args.steps} "
            f"solves, {accepted} accepted); judged every solve {ms_every:.3f} ms, guess residual "
            f"{st_every['learned_residual']:.3e}")
        log(f"bench: learned correction {ms_corr:.3f} ms/solve, {st_corr['cycles']} cycles, kept {corr_kept} of "
            f"{args.steps}, ||r - alpha A c|| / ||r|| = {st_corr['correction_residual']:.3e}; forced: "
            f"{st_corr_forced['cycles']} cycles")
        g.attach_learned(None)

    # the roofline kernel inside the headline solves: HIP events around every
    # launch of the FCG update fused into the FP32 level-0 pre-smoothing (two
    # red-black sweeps + residual + restriction; every iteration after the
    # first), over a second timed pass of the same rotating solves
    log("bench: in-solve launches of the roofline kernel (HIP events, second pass)")
    g.probe_kernel(6, capacity=24 * (args.steps + args.warmup))
    if learned.get("attached"):
        g.attach_learned(net, meta["level"])
        ms_probe, _, _, _ = backoff(1)
        g.attach_learned(None)
    else:
        ms_probe, _ = run(PREC_MIXED, args.steps, args.warmup, accel=1, lagged=1)
    launches, ev_total_ms, k_total_ms, kbytes = g.probe_result()
    g.probe_kernel(-1, 0)
    kms = ev_total_ms / max(launches, 1)       # HIP events around each launch (timing-only events)
    span_ms = k_total_ms / max(launches, 1)    # its workgroups' first start to last end (device clock)
    achieved = kbytes / (kms * 1e-3) / 1e9
    # level-0 kernels relaunched back to back by the timing hook (HIP events)
    kernels = {}
    for which, name in [(0, "fp64_pre_smooth_restrict"), (1, "fp32_pre_smooth_restrict"),
                        (2, "fp32_post_smooth_fp64_outer"), (3, "fp64_post_smooth_norm")]:
        ms_k, b_k = g.time_kernel(which, iters=50)
        kernels[name] = {"us": round(ms_k * 1e3, 2), "bytes": b_k, "GBps": round(b_k / (ms_k * 1e-3) / 1e9, 1)}
    kernels["fp32_pre_smooth_restrict"]["traffic"] = round(PRE2D_FP32_TRAFFIC_BYTES)
    kernels["note"] = ("back-to-back relaunches of one kernel over the same arrays (cache-warm); the roofline "
                       "kernel is timed inside the solves instead")

    # U-Net training throughput on the matrix cores (level-1 grid, batch 4)
    unet = None
    if not args.no_unet:
        log("bench: U-Net training step (256x512, batch 4)")
        unet = bench_unet()

    momentum = None
    if not args.no_momentum:
        log("bench: momentum predictor at the 1024x512 dam-break shape")
        momentum = bench_momentum()
        log(f"bench: momentum {momentum['ms_per_step']} ms/step, {momentum['GBps']} GB/s")

    interface = None
    if not args.no_interface:
        log("bench: interface maintenance at the 1024x512 dam-break shape")
        interface = bench_interface()
        log(f"bench: uccd rk3 {interface['uccd_rk3_step_ms']} ms, mpls x5 {interface['mpls_5_steps_ms']} ms, "
            f"redistance {interface['redistance_rk3_step_ms']} ms")

    log("bench: streaming ceilings (device copy, three-read one-write pass)")
    ceiling = bench_hbm_ceiling(torch)
    log(f"bench: copy {ceiling['copy_GBps']} GB/s, 3R1W {ceiling['three_read_one_write_GBps']} GB/s")

    weno = None
    if not args.no_weno:
        log("bench: WENO5 + TVD-RK3 level-set step at the 1024x512 dam-break shape")
        weno = bench_weno()
        log(f"bench: weno rk3 {weno['ms_per_step']} ms/step, {weno['GBps']} GB/s")

    snapshot = None
    if not args.no_snapshot:
        log("bench: VTK snapshot of the 512x1024 grid")
        snapshot = bench_snapshot(torch)
        log(f"bench: snapshot {snapshot['snapshot_ms']} ms ({snapshot['snapshot_MBps']} MB/s), "
            f"pack {snapshot['pack_scalar_GBps']} GB/s")

    out = {
        "metric": METRIC,
        "value": round(ms_ml, 4),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_ml, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 V-cycle + fp64 defect correction; bf16 U-Net",
        "data": "synthetic (seeded bubble density 1000:1, random smooth divergence RHS)",
        "config": {"workload": "bubble_2d 512x1024 variable-density pressure Poisson, all-Neumann",
                   "rel_tol": tol, "cycle": "V(2,2) red-black GS", "levels": g.level_num,
                   "outer": "FCG(1) defect correction",
                   "learned": "U-Net guess, FP64-judged, back-off after rejections",
                   "convergence_check": "lagged (residual read from one iteration before the last count)"},
        "fp64_mg": {"ms_per_solve": round(ms64, 4), "cycles": st64["cycles"],
                    "rel_residual": st64["rel_residual"], "accel": "stationary V-cycles"},
        "fp64_mg_fcg": {"ms_per_solve": round(ms64k, 4), "cycles": st64k["cycles"],
                        "rel_residual": st64k["rel_residual"]},
        "fp64_mg_lagged": {"ms_per_solve": round(ms64L, 4), "cycles": st64L["cycles"],
                           "rel_residual": st64L["rel_residual"], "accel": "stationary V-cycles",
                           "lagged_check": True, "cycles_per_solve": st64L["cycles_per_solve"]},
        "mixed_lagged": {"ms_per_solve": round(ms_mxL, 4), "cycles": st_mxL["cycles"],
                         "rel_residual": st_mxL["rel_residual"], "rel_l2_vs_fp64": errL, "lagged_check": True,
                         "cycles_per_solve": st_mxL["cycles_per_solve"]},
        "mixed": {"ms_per_solve": round(ms_mx, 4), "cycles": st_mx["cycles"],
                  "rel_residual": st_mx["rel_residual"], "rel_l2_vs_fp64": err,
                  "accel": "flexible CG (FCG(1)), FP64 vectors", "learned": False},
        "mixed_learned": learned,
        "speedup_vs_fp64": round(ms64 / ms_ml, 3),
        "speedup_vs_fp64_lagged": round(ms64L / ms_ml, 3),
        "speedup_mixed_vs_fp64": round(ms64 / ms_mx, 3),
        "speedup_mixed_vs_fp64_fcg": round(ms64k / ms_mx, 3),
        "kernels": kernels,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(FCG_PRE2D_TRAFFIC_BYTES),
                     "kernel": "k_pre2d_fcg<true,32,16> level 0 (FCG update of x and r + 2 RB-GS sweeps + "
                               "residual + restriction)",
                     "timing": f"HIP events (timing-only, no system fence) around each of its {launches} "
                               f"launches inside the headline solves (second timed pass, {round(ms_probe, 4)} "
                               f"ms/solve) on the solver stream; us_per_launch_span: first workgroup start to "
                               f"last workgroup end on the device wall clock",
                     "us_per_launch_span": round(span_ms * 1e3, 3),
                     "note": "largest HBM-facing kernel; the largest by time is the one-workgroup coarse-level "
                             "cycle k_block_vcycle<float,2> (on-chip, one CU)",
                     "bytes_per_launch": kbytes, "us_per_launch": round(kms * 1e3, 3),
                     "launches": launches, "measured_ceiling": ceiling},
        "unet": unet,
        "momentum": momentum,
        "interface": interface,
        "weno_rk3": weno,
        "snapshot": snapshot,
    }
    if not args.no_3d:
        out["plane_3d_256"] = bench_3d(args)
    if not args.no_learned3d:
        log("bench: 3-D learned guess on held-out 128^3 bubbles")
        out["learned_3d_128"] = bench_learned3d()
        log("bench: dam-break learned guess on held-out steps")
        out["learned_dambreak"] = bench_learned_dambreak()
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle port, 1 core)")
        out["cpu_baseline"] = cpu_baseline(w, rhs_host, args.cpu_seconds)
        if weno is not None:
            out["cpu_baseline"]["weno_rk3"] = cpu_weno_baseline()
        if snapshot is not None:
            out["cpu_baseline"]["snapshot_writer"] = cpu_snapshot_baseline(snapshot)
        if momentum is not None:
            out["cpu_baseline"]["momentum"] = cpu_momentum_baseline()
        if interface is not None:
            out["cpu_baseline"]["interface"] = cpu_interface_baseline()
    if snapshot is not None:
        snapshot.pop("_fields")
        snapshot.pop("_file")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
