"""Training data, training loop and weights of the learned pressure guess.

The U-Net (``scihpc_amd.UNet``; no SciHPC counterpart) is trained on
pressure systems that the GPU solver itself writes and solves:

  --data bubble    randomised variable-density bubbles in the 1 x 2 box
                   (512 x 1024, 1000:1; centre, radius and the divergence
                   modes drawn from a seeded generator);
  --data bubble3d  random spheres of radius 0.2 in the unit cube (128^3,
                   1000:1) for the 3-D network;
  --data dambreak  snapshots of the 1024 x 512 dam break run by
                   drivers/dambreak_demo (SciHPC's benchmarks_2d_dambreak.cpp
                   loop through its own API), read back with read_vtk: the
                   projection's right-hand side div(u*)/dt and the density of
                   every 10th step.  The last --holdout snapshots are held out.

Each system is solved in FP64 to 1e-9, and ``Multigrid.learned_features``
takes the network's input features and its target (the solution,
block-averaged onto the network's multigrid level and divided by the RMS of
b) on the device.  Training runs on the GPU with the library's own BF16
forward/backward kernels and Adam; mirroring along the axes is an exact
symmetry of the all-Neumann box problem and is used as augmentation.

The network runs on a coarse multigrid level (level 3, 64 x 128, for the
512 x 1024 bubble; level 2, 32^3, for 128^3): there its four pooling levels
see the whole box, which the long-range pressure response needs.  The loss
is the A(rho) energy norm of the error, optionally after damped-Jacobi
sweeps of it (--smooth): a guess fitted in the mean square carries
pixel-scale noise whose energy inside the light fluid (1000x larger
coefficients) exceeds the solution's own.

Command lines (one GPU; the shipped weights were made this way):
    python -m scihpc_amd.learned --data bubble --level 3 --samples 256 --steps 4000 --smooth 2
    python -m scihpc_amd.learned --data bubble3d --level 2 --samples 96 --steps 3000 --smooth 2
    python -m scihpc_amd.learned --data dambreak --run-steps 2000 --cfl 0.1 --holdout 30 --smooth 2
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

from . import PREC_FP64, PREC_MIXED, Multigrid, UNet, read_vtk
from . import workloads as W

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "unet_bubble_512x1024_level3.npz")
TRAIN_SEED = 7_000_000  # seeds of training problems; evaluation problems use EVAL_SEED
EVAL_SEED = 9_000_000


def bubble_configs(count, seed):
    """Random single-bubble problems in the 1 x 2 box."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        r = float(rng.uniform(0.12, 0.3))
        out.append(dict(center=(float(rng.uniform(r + 0.05, 0.95 - r)), float(rng.uniform(r + 0.05, 1.95 - r))),
                        radius=r, seed=int(seed + 1 + k)))
    return out


def make_problem(cfg, nx=512, ny=1024):
    if len(cfg["center"]) == 3:
        return W.bubble_3d(nx, center=cfg["center"], radius=cfg["radius"], seed=cfg["seed"])
    return W.bubble_2d(nx, ny, center=cfg["center"], radius=cfg["radius"], seed=cfg["seed"])


def bubble3d_configs(count, seed, radius=0.2):
    """Random spheres of radius 0.2 in the unit cube (the 3-D bubble of the
    128^3 problems): centre uniform with the sphere 0.05 clear of the walls."""
    rng = np.random.default_rng(seed)
    lo, hi = radius + 0.05, 0.95 - radius
    return [dict(center=tuple(float(v) for v in rng.uniform(lo, hi, 3)), radius=radius, seed=int(seed + 1 + k))
            for k in range(count)]


DAMBREAK_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "drivers", "dambreak_demo")


def dambreak_run(out_dir, nx=1024, ny=512, steps=2000, cfl=0.25, every=10, log=None):
    """The dam break of benchmarks_2d_dambreak.cpp:18-131 through the
    reference-API driver (drivers/dambreak_demo): box 4 x 2, water column
    1 x 2 at the left wall started from its signed distance, air/water
    (rho 1000 : 1.226, mu 1.137e-3 : 1.78e-5), gravity, no-slip walls,
    dt = cfl h, snapshots <out_dir>/db<id>.vtk every `every` steps (id = step
    / every; id 0 is the initial state).  Returns the driver's key/value
    results; its progress lines go to stderr as they come."""
    os.makedirs(out_dir, exist_ok=True)
    if not os.path.exists(DAMBREAK_DRIVER):
        raise RuntimeError("drivers/dambreak_demo is not built (scihpc_amd.build)")
    cmd = [DAMBREAK_DRIVER, str(nx), str(ny), "4", "2", "1", "2", str(steps), str(cfl), "distance",
           os.path.join(out_dir, "db"), str(every)]
    if log:
        log("dam break: " + " ".join(cmd))
    res = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
    out = {}
    for line in res.stdout.splitlines():
        k, v = line.split()
        out[k] = float(v)
    # status 3: the flow stopped being finite at step stopped_at; the
    # snapshots written before it are whole and are kept
    if res.returncode != 0 and not (res.returncode == 3 and "stopped_at" in out):
        raise RuntimeError(f"dambreak_demo failed with status {res.returncode}")
    return out


def dambreak_snapshots(out_dir):
    """Sorted [(id, path)] of the snapshots db<id>.vtk in out_dir."""
    found = []
    for name in os.listdir(out_dir):
        if name.startswith("db") and name.endswith(".vtk") and name[2:-4].isdigit():
            found.append((int(name[2:-4]), os.path.join(out_dir, name)))
    return sorted(found)


def dambreak_problem(path):
    """The pressure system of one snapshot, read back with read_vtk:
    A(rho) p = b with b = div(u*) / dt (the projection right-hand side,
    projection_method.cpp:546-568), its mean removed (all-Neumann
    compatibility), and rho = the snapshot's density divided by its
    minimum (air = 1, water = 815.7), the convention of the bubble
    problems.  Arrays in the multigrid layout (x fastest)."""
    snap = read_vtk(path)
    rho = snap["density"][0]
    b = snap["rhs"][0].astype(np.float64)
    return dict(rho=(rho / rho.min()).ravel(), rhs=(b - b.mean()).ravel(), pressure=snap["pressure"][0].ravel(),
                phi=snap["phi"][0].ravel(), dims=snap["dims"], spacing=snap["spacing"])


DAMBREAK_FIXTURE = os.path.join(WEIGHTS_DIR, "dambreak_heldout_1024x512.npz")


def dambreak_fixture(snap_dir, every, steps, out=DAMBREAK_FIXTURE):
    """Held-out dam-break pressure systems for the tests and the bench: the
    right-hand side and density of the snapshots of the given steps
    (FP32), from a run writing a snapshot every `every` steps."""
    arrs = {}
    for st in steps:
        snap = read_vtk(os.path.join(snap_dir, f"db{st // every}.vtk"))
        arrs[f"rhs_{st}"] = snap["rhs"][0].astype(np.float32)
        arrs[f"density_{st}"] = snap["density"][0].astype(np.float32)
    np.savez_compressed(out, **arrs)
    return out


def dambreak_fixture_problems(path=DAMBREAK_FIXTURE):
    """{step: problem} of the held-out fixture, as dambreak_problem gives them."""
    out = {}
    with np.load(path, allow_pickle=False) as z:
        for key in z.files:
            if key.startswith("rhs_"):
                st = int(key[4:])
                b = z[key].astype(np.float64)
                rho = z[f"density_{st}"].astype(np.float64)
                out[st] = dict(rho=(rho / rho.min()).ravel(), rhs=(b - b.mean()).ravel())
    return out


def solver_samples(mg: Multigrid, configs, level=1, rel_tol=1e-9, log=None):
    """(features [n, H*W, 2], targets [n, H*W], faces [n, H*W, 2]) written
    by the GPU solver.  configs: bubble configurations, or problems (dicts
    with "rho" and "rhs") such as dambreak_problem's."""
    nx, ny = mg.shapes[0][0], mg.shapes[0][1]
    feats, tgts, faces = [], [], []
    for k, cfg in enumerate(configs):
        w = cfg if "rhs" in cfg else make_problem(cfg, nx, ny)
        mg.set_density(w["rho"])
        mg.set_rhs(w["rhs"])
        st = mg.solve(rel_tol=rel_tol, max_cycles=100, precision=PREC_FP64, accel=1)
        if not st["rel_residual"] <= 100 * rel_tol:  # a target 100x looser than asked is still exact enough
            raise RuntimeError(f"training solve {k} did not converge: {st}")
        f, t, c = mg.learned_features(level, with_target=True, with_faces=True)
        feats.append(f)
        tgts.append(t)
        faces.append(c)
        if log and (k + 1) % 16 == 0:
            log(f"samples: {k + 1}/{len(configs)}")
    return np.stack(feats), np.stack(tgts), np.stack(faces)


def train(net: UNet, feat, tgt, steps=2000, batch=4, lr=1e-3, lr_min=1e-5, seed=0, log=None,
          time_limit=None, log_every=200, faces=None, smooth=0):
    """Adam with a cosine learning-rate schedule; returns the loss history.
    With faces the loss is the relative A(rho) energy norm of the error
    (after `smooth` damped-Jacobi sweeps of it)."""
    net.out_scale = float(np.sqrt(np.mean(tgt.astype(np.float64) ** 2)))
    net.set_samples(feat, tgt, faces)
    net.set_loss_smoothing(smooth if faces is not None else 0)
    rng = np.random.default_rng(seed)
    n = feat.shape[0]
    hist = []
    t0 = time.perf_counter()
    for s in range(steps):
        frac = s / max(1, steps - 1)
        rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * frac))
        idx = rng.integers(0, n, size=batch)
        flips = rng.integers(0, 2 ** net.ndim, size=batch)
        want = (s % log_every == 0) or s == steps - 1
        loss = net.train_step(idx, rate, flips, want_loss=want)
        if want:
            hist.append((s, loss))
            if log:
                log(f"step {s}: loss {loss:.4e} lr {rate:.2e} ({time.perf_counter() - t0:.1f} s)")
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            break
    net.sync()
    return hist


def save_weights(path, net: UNet, level, meta=None):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=net.params(), out_scale=np.float32(net.out_scale), level=np.int32(level),
             height=np.int32(net.H), width=np.int32(net.W), depth=np.int32(net.D if net.ndim == 3 else 0),
             meta=np.array(json.dumps(meta or {})))


def load_weights(path):
    with np.load(path, allow_pickle=False) as z:
        return dict(params=z["params"], out_scale=float(z["out_scale"]), level=int(z["level"]),
                    height=int(z["height"]), width=int(z["width"]),
                    depth=int(z["depth"]) if "depth" in z.files else 0, meta=json.loads(str(z["meta"])))


def load_net(path=DEFAULT_WEIGHTS):
    w = load_weights(path)
    net = UNet(w["height"], w["width"], 1, depth=w["depth"] or None)
    net.set_params(w["params"])
    net.out_scale = w["out_scale"]
    return net, w


def held_out_gain(mg: Multigrid, configs, rel_tol=1e-8, precision=PREC_MIXED, accel=1, reps=1):
    """Per problem: FCG iterations and ms with and without the learned guess."""
    nx, ny = mg.shapes[0][0], mg.shapes[0][1]
    rows = []
    for cfg in configs:
        w = cfg if "rhs" in cfg else make_problem(cfg, nx, ny)
        mg.set_density(w["rho"])
        mg.set_rhs(w["rhs"])
        res = {}
        for learned in (0, 3, 2):  # 3: judged on every solve (no back-off); 2: forced
            best = None
            for _ in range(reps):
                st = mg.solve(rel_tol=rel_tol, max_cycles=100, precision=precision, accel=accel,
                              use_learned=learned)
                best = st if best is None or st["ms"] < best["ms"] else best
            res[learned] = best
        rows.append(dict(cfg=cfg, plain=res[0], learned=res[3], forced=res[2]))
    return rows


def guess_quality(mg: Multigrid, problems, rel_tol=1e-10):
    """Per problem: the attached network's guess (the solver's own pipeline,
    scaled by its energy-optimal alpha) against the FP64 solution:
    relative energy-norm error ||x - g||_A / ||x||_A, relative L2 error and
    relative residual ||b - A g|| / ||b||."""
    nx, ny = mg.shapes[0][0], mg.shapes[0][1]
    rows = []
    for cfg in problems:
        w = cfg if "rhs" in cfg else make_problem(cfg, nx, ny)
        mg.set_density(w["rho"])
        mg.set_rhs(w["rhs"])
        mg.solve(rel_tol=rel_tol, max_cycles=200, precision=PREC_FP64, accel=1)
        x = mg.sol(0).copy()
        st = mg.solve(rel_tol=rel_tol, max_cycles=0, precision=PREC_FP64, use_learned=2)
        g = mg.sol(0).copy()
        e = x - g
        ax, ae = mg.apply(x), mg.apply(e)
        rows.append(dict(energy=float(np.sqrt(abs(e @ ae) / abs(x @ ax))),
                         l2=float(np.linalg.norm(e) / np.linalg.norm(x)), residual=st["rel_residual"]))
    return rows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--data", choices=["bubble", "bubble3d", "dambreak"], default="bubble")
    ap.add_argument("--samples", type=int, default=256, help="bubble problems to train on")
    ap.add_argument("--snapshots", default=None, help="dam break: directory of db<id>.vtk snapshots")
    ap.add_argument("--run-steps", type=int, default=0,
                    help="dam break: run the 1024x512 dam break for this many steps first (0: reuse snapshots)")
    ap.add_argument("--every", type=int, default=10, help="dam break: snapshot every k steps")
    ap.add_argument("--cfl", type=float, default=0.1,
                    help="dam break: largest dt / h (find_dt's stability bound caps it further)")
    ap.add_argument("--holdout", type=int, default=40, help="dam break: the last k snapshots are held out")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--level", type=int, default=3)
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--eval", type=int, default=8, help="held-out problems to evaluate")
    ap.add_argument("--mse", action="store_true", help="mean-squared-error loss instead of the energy norm")
    ap.add_argument("--smooth", type=int, default=0,
                    help="damped-Jacobi sweeps of the error inside the energy loss")
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None, help="write the run's summary (JSON) here")
    args = ap.parse_args()

    def log(m):
        print(m, file=sys.stderr, flush=True)

    summary = {}
    if args.data == "dambreak":
        nx, ny, lx, ly = 1024, 512, 4.0, 2.0
        snap_dir = args.snapshots or os.path.join(os.environ.get("TMPDIR", "/tmp"), "scihpc_dambreak")
        if args.run_steps > 0:
            t0 = time.perf_counter()
            summary["run"] = dambreak_run(snap_dir, nx, ny, args.run_steps, args.cfl, args.every, log=log)
            summary["run"]["wall_seconds"] = round(time.perf_counter() - t0, 1)
            log(f"dam break: {json.dumps(summary['run'])}")
        snaps = [(i, f) for i, f in dambreak_snapshots(snap_dir) if i > 0]  # id 0: before the first step
        if "run" in summary and "stopped_at" in summary["run"]:
            # the flow ran away before it stopped: snapshots of the last 50
            # steps before the stop are left out
            last = (int(summary["run"]["stopped_at"]) - 50) // args.every
            snaps = [(i, f) for i, f in snaps if i < last]
        if len(snaps) <= args.holdout:
            raise SystemExit(f"{len(snaps)} snapshots in {snap_dir}: not enough for {args.holdout} held out")
        t0 = time.perf_counter()
        problems = [dambreak_problem(f) for _, f in snaps]
        log(f"read {len(problems)} snapshots with read_vtk in {time.perf_counter() - t0:.1f} s")
        assert all(p["dims"] == (nx + 1, ny + 1, 1) for p in problems)
        train_set, eval_set = problems[:-args.holdout], problems[-args.holdout:]
        summary["snapshots"] = dict(dir=snap_dir, count=len(snaps), train_ids=[snaps[0][0], snaps[-args.holdout - 1][0]],
                                    held_out_ids=[snaps[-args.holdout][0], snaps[-1][0]], every=args.every)
        out = args.out or os.path.join(WEIGHTS_DIR, f"unet_dambreak_{nx}x{ny}_level{args.level}.npz")
    elif args.data == "bubble3d":
        nx = ny = 128
        train_set = bubble3d_configs(args.samples, TRAIN_SEED)
        eval_set = bubble3d_configs(args.eval, EVAL_SEED)
        out = args.out or os.path.join(WEIGHTS_DIR, f"unet_bubble3d_128_level{args.level}.npz")
    else:
        nx, ny, lx, ly = 512, 1024, 1.0, 2.0
        train_set = bubble_configs(args.samples, TRAIN_SEED)
        eval_set = bubble_configs(args.eval, EVAL_SEED)
        out = args.out or DEFAULT_WEIGHTS
    if args.data == "bubble3d":
        mg = Multigrid(3, nx, nx, nx, 1 / nx, 1 / nx, 1 / nx, "neumann")
    else:
        mg = Multigrid(2, nx, ny, 1, lx / nx, ly / ny, 1.0, "neumann")
    H, Wd, Dd = mg.shapes[args.level][1], mg.shapes[args.level][0], mg.shapes[args.level][2]
    depth = Dd if args.data == "bubble3d" else None
    t0 = time.perf_counter()
    feat, tgt, faces = solver_samples(mg, train_set, args.level, log=log)
    log(f"{len(train_set)} samples in {time.perf_counter() - t0:.1f} s")
    net = UNet(H, Wd, args.batch, seed=1, depth=depth)
    t0 = time.perf_counter()
    hist = train(net, feat, tgt, args.steps, args.batch, args.lr, log=log, time_limit=args.time_limit,
                 faces=None if args.mse else faces, smooth=args.smooth)
    t_train = time.perf_counter() - t0
    log(f"trained {hist[-1][0] + 1} steps in {t_train:.1f} s")
    meta = dict(data=args.data, samples=len(train_set), steps=hist[-1][0] + 1, batch=args.batch, lr=args.lr,
                loss="mse" if args.mse else ("energy" if not args.smooth else f"energy after {args.smooth} Jacobi sweeps"),
                train_seconds=round(t_train, 1), loss_first=hist[0][1], loss_last=hist[-1][1],
                grid=[nx, ny])
    if args.data != "dambreak":
        meta["train_seed"] = TRAIN_SEED
    else:
        meta["snapshots"] = summary["snapshots"]
    save_weights(out, net, args.level, meta)
    log(f"weights -> {out}")
    summary["training"] = meta
    if args.eval:
        infer = UNet(H, Wd, 1, depth=depth)
        infer.set_params(net.params())
        infer.out_scale = net.out_scale
        mg.attach_learned(infer, args.level)
        evals = eval_set[:args.eval] if args.data != "dambreak" else eval_set
        q = guess_quality(mg, evals)
        rows = held_out_gain(mg, evals)
        for r, g in zip(rows, q):
            log(f"held-out: plain {r['plain']['cycles']} it {r['plain']['ms']:.3f} ms, "
                f"learned {r['learned']['cycles']} it {r['learned']['ms']:.3f} ms, "
                f"forced {r['forced']['cycles']} it {r['forced']['ms']:.3f} ms "
                f"(guess energy error {g['energy']:.3e}, L2 {g['l2']:.3e}, residual {g['residual']:.3e}; "
                f"accepted {r['learned']['learned_accepted']})")
        summary["held_out"] = dict(
            problems=len(rows),
            plain_cycles=float(np.mean([r["plain"]["cycles"] for r in rows])),
            learned_cycles=float(np.mean([r["learned"]["cycles"] for r in rows])),
            plain_ms=float(np.mean([r["plain"]["ms"] for r in rows])),
            learned_ms=float(np.mean([r["learned"]["ms"] for r in rows])),
            forced_cycles=float(np.mean([r["forced"]["cycles"] for r in rows])),
            forced_ms=float(np.mean([r["forced"]["ms"] for r in rows])),
            accepted=int(sum(r["learned"]["learned_accepted"] for r in rows)),
            guess_energy_error=float(np.median([g["energy"] for g in q])),
            guess_l2_error=float(np.median([g["l2"] for g in q])),
            guess_residual=float(np.median([g["residual"] for g in q])))
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(summary, fh, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
