"""Training data, training loop and weights of the learned pressure guess.

The U-Net (``scihpc_amd.UNet``; no SciHPC counterpart) is trained on
snapshots that the GPU multigrid itself writes: randomised variable-density
bubble problems (centre, radius and the seed of the divergence modes drawn
from a seeded generator; 1000:1 density) are solved in FP64 to 1e-9 relative
residual, and ``Multigrid.learned_features`` takes the network's input
features and its target (the solution, block-averaged onto the network's
multigrid level and divided by the RMS of b) on the device.  Training runs
on the GPU with the library's own BF16 forward/backward kernels and Adam;
mirroring x and/or y is an exact symmetry of the all-Neumann box problem
and is used as augmentation.

The network runs on multigrid level 3 (64 x 128 for the 512 x 1024 bubble):
there its four pooling levels see the whole box, which the long-range
pressure response needs; on finer levels held-out errors stay near 100 %.
The loss is the A(rho) energy norm of the error: a guess fitted in the mean
square carries pixel-scale noise whose energy inside the light fluid (1000x
larger coefficients) exceeds the solution's own.

Command line (one GPU; the shipped weights were made this way):
    python -m scihpc_amd.learned --level 3 --samples 256 --steps 4000
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import PREC_FP64, PREC_MIXED, Multigrid, UNet
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
    return W.bubble_2d(nx, ny, center=cfg["center"], radius=cfg["radius"], seed=cfg["seed"])


def solver_samples(mg: Multigrid, configs, level=1, rel_tol=1e-9, log=None):
    """(features [n, H*W, 2], targets [n, H*W], faces [n, H*W, 2]) written
    by the GPU solver."""
    nx, ny = mg.shapes[0][0], mg.shapes[0][1]
    feats, tgts, faces = [], [], []
    for k, cfg in enumerate(configs):
        w = make_problem(cfg, nx, ny)
        mg.set_density(w["rho"])
        mg.set_rhs(w["rhs"])
        st = mg.solve(rel_tol=rel_tol, max_cycles=100, precision=PREC_FP64, accel=1)
        if not st["converged"]:
            raise RuntimeError(f"training solve {k} did not converge: {st}")
        f, t, c = mg.learned_features(level, with_target=True, with_faces=True)
        feats.append(f)
        tgts.append(t)
        faces.append(c)
        if log and (k + 1) % 16 == 0:
            log(f"samples: {k + 1}/{len(configs)}")
    return np.stack(feats), np.stack(tgts), np.stack(faces)


def train(net: UNet, feat, tgt, steps=2000, batch=4, lr=1e-3, lr_min=1e-5, seed=0, log=None,
          time_limit=None, log_every=200, faces=None):
    """Adam with a cosine learning-rate schedule; returns the loss history.
    With faces the loss is the relative A(rho) energy norm of the error."""
    net.out_scale = float(np.sqrt(np.mean(tgt.astype(np.float64) ** 2)))
    net.set_samples(feat, tgt, faces)
    rng = np.random.default_rng(seed)
    n = feat.shape[0]
    hist = []
    t0 = time.perf_counter()
    for s in range(steps):
        frac = s / max(1, steps - 1)
        rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * frac))
        idx = rng.integers(0, n, size=batch)
        flips = rng.integers(0, 4, size=batch)
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
             height=np.int32(net.H), width=np.int32(net.W), meta=np.array(json.dumps(meta or {})))


def load_weights(path):
    with np.load(path, allow_pickle=False) as z:
        return dict(params=z["params"], out_scale=float(z["out_scale"]), level=int(z["level"]),
                    height=int(z["height"]), width=int(z["width"]), meta=json.loads(str(z["meta"])))


def load_net(path=DEFAULT_WEIGHTS):
    w = load_weights(path)
    net = UNet(w["height"], w["width"], 1)
    net.set_params(w["params"])
    net.out_scale = w["out_scale"]
    return net, w


def held_out_gain(mg: Multigrid, configs, rel_tol=1e-8, precision=PREC_MIXED, accel=1, reps=1):
    """Per problem: FCG iterations and ms with and without the learned guess."""
    nx, ny = mg.shapes[0][0], mg.shapes[0][1]
    rows = []
    for cfg in configs:
        w = make_problem(cfg, nx, ny)
        mg.set_density(w["rho"])
        mg.set_rhs(w["rhs"])
        res = {}
        for learned in (0, 3):  # 3: judged on every solve (no back-off)
            best = None
            for _ in range(reps):
                st = mg.solve(rel_tol=rel_tol, max_cycles=100, precision=precision, accel=accel,
                              use_learned=learned)
                best = st if best is None or st["ms"] < best["ms"] else best
            res[learned] = best
        rows.append(dict(cfg=cfg, plain=res[0], learned=res[3]))
    return rows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--samples", type=int, default=256)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--level", type=int, default=3)
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--eval", type=int, default=8, help="held-out problems to evaluate")
    ap.add_argument("--mse", action="store_true", help="mean-squared-error loss instead of the energy norm")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args()

    def log(m):
        print(m, file=sys.stderr, flush=True)

    mg = Multigrid(2, 512, 1024, 1, 1 / 512, 2 / 1024, 1.0, "neumann")
    H, Wd = mg.shapes[args.level][1], mg.shapes[args.level][0]
    t0 = time.perf_counter()
    feat, tgt, faces = solver_samples(mg, bubble_configs(args.samples, TRAIN_SEED), args.level, log=log)
    log(f"{args.samples} samples in {time.perf_counter() - t0:.1f} s")
    net = UNet(H, Wd, args.batch, seed=1)
    t0 = time.perf_counter()
    hist = train(net, feat, tgt, args.steps, args.batch, args.lr, log=log, time_limit=args.time_limit,
                 faces=None if args.mse else faces)
    t_train = time.perf_counter() - t0
    log(f"trained {hist[-1][0] + 1} steps in {t_train:.1f} s")
    meta = dict(samples=args.samples, steps=hist[-1][0] + 1, batch=args.batch, lr=args.lr,
                loss="mse" if args.mse else "energy",
                train_seconds=round(t_train, 1), loss_first=hist[0][1], loss_last=hist[-1][1],
                train_seed=TRAIN_SEED, grid=[512, 1024])
    save_weights(args.out, net, args.level, meta)
    log(f"weights -> {args.out}")
    if args.eval:
        infer = UNet(H, Wd, 1)
        infer.set_params(net.params())
        infer.out_scale = net.out_scale
        mg.attach_learned(infer, args.level)
        rows = held_out_gain(mg, bubble_configs(args.eval, EVAL_SEED))
        for r in rows:
            log(f"held-out r={r['cfg']['radius']:.3f}: plain {r['plain']['cycles']} it {r['plain']['ms']:.3f} ms, "
                f"learned {r['learned']['cycles']} it {r['learned']['ms']:.3f} ms "
                f"(first-step residual {r['learned']['learned_residual']:.3e}, "
                f"accepted {r['learned']['learned_accepted']})")
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
