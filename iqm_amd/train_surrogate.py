"""Trains the phase-congruency surrogate on the device and evaluates it.

Training data: `--patches` seeded size x size gray 1/f^2 patches generated on
the device (half distorted: contrast, noise, 16-level quantisation; with
`--scale k` made at k*size and box-downscaled by FSIM as large inputs are), targets
the exact PC maps of the library's FSIM path; Adam on minibatches in a seeded
order, one checkpoint per epoch (weights, Adam moments, step, cursor), and
`--resume` continues from a checkpoint.  `--loss-eps e` (e >= 0) weights each
pixel's squared error by max(PC of the pair's two images) + e, the weight FSIM
gives that pixel's similarity; the committed network was trained with e = 0.
Evaluation: FSIM/FSIMc of seeded 1920x1080 RGB pairs with exact PC and with
network PC, and the spot-check error the fallback uses.

  python -m iqm_amd.train_surrogate --scale 4 --loss-eps 0 --out pc_surrogate.iqmsur
  python -m iqm_amd.train_surrogate --eval-only --resume pc_surrogate.iqmsur
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import iqm_amd
from iqm_amd import synth


def batch_order(n_batches: int, seed: int, epoch: int) -> np.ndarray:
    return np.random.default_rng([seed, epoch]).permutation(n_batches)


def train(ctx, sur, args) -> list:
    P, S, B = args.patches, args.size, args.batch
    if P % B or B % 2:
        raise SystemExit("patches must be a multiple of an even batch size")
    px = S * S
    dx, dp = ctx.alloc(4 * P * px), ctx.alloc(4 * P * px)
    t0 = time.perf_counter()
    sur.make_patches(P, S, args.seed, dx, dp, scale=args.scale)
    print(f"[train] {P} patches of {S}x{S} (made at {S * args.scale}^2) in {time.perf_counter() - t0:.2f} s",
          flush=True)
    nb = P // B
    step0, _ = sur.state()
    hist = []
    for epoch in range(step0 // nb, args.epochs):
        order = batch_order(nb, args.seed, epoch)
        first = step0 - epoch * nb if epoch == step0 // nb else 0
        t0 = time.perf_counter()
        losses = []
        for k in range(max(first, 0), nb):
            off = int(order[k]) * B * px * 4
            lr = args.lr * (0.5 ** (epoch // max(1, args.lr_halve)))
            losses.append(sur.train_step(dx + off, dp + off, B, S, S, lr))
        dt = time.perf_counter() - t0
        m = float(np.mean(losses)) if losses else float("nan")
        hist.append({"epoch": epoch, "mse": m, "seconds": dt, "patches_per_s": len(losses) * B / dt if dt else 0})
        print(f"[train] epoch {epoch}: mse {m:.5e}  {dt:.2f} s  ({hist[-1]['patches_per_s']:.0f} patches/s)",
              flush=True)
        if args.out:
            with open(args.out, "wb") as f:
                f.write(sur.save())
    ctx.free(dx)
    ctx.free(dp)
    return hist


def evaluate(ctx, sur, n_pairs: int, seed: int, spot_tol: float) -> dict:
    W, H = 1920, 1080
    fsim = iqm_amd.FSIM()
    rows = []
    for i in range(n_pairs):
        img, ref = synth.rgb_pair(W, H, seed=seed + i)
        di, dr = ctx.alloc(img.nbytes), ctx.alloc(ref.nbytes)
        ctx.upload(di, img)
        ctx.upload(dr, ref)
        ex_fs, ex_fc = fsim.compute_batch_dev(ctx, di, dr, W, H, 3, img.nbytes, 1)
        fs, fc, err, back = sur.fsim_batch_dev(di, dr, W, H, 3, img.nbytes, 1, spot_every=1, spot_tol=spot_tol)
        fs0, fc0, _, _ = sur.fsim_batch_dev(di, dr, W, H, 3, img.nbytes, 1, spot_every=0)
        ctx.free(di)
        ctx.free(dr)
        rows.append((ex_fs[0], ex_fc[0], fs0[0], fc0[0], err, back))
        print(f"[eval] pair {seed + i}: exact {ex_fs[0]:.5f}/{ex_fc[0]:.5f} net {fs0[0]:.5f}/{fc0[0]:.5f} "
              f"spot {err:.4f}{' fallback' if back else ''}", flush=True)
    a = np.array([r[:5] for r in rows], np.float64)
    return {"pairs": n_pairs, "seeds": [seed, seed + n_pairs - 1], "size": [W, H],
            "mean_abs_dfsim": float(np.mean(np.abs(a[:, 2] - a[:, 0]))),
            "max_abs_dfsim": float(np.max(np.abs(a[:, 2] - a[:, 0]))),
            "mean_abs_dfsimc": float(np.mean(np.abs(a[:, 3] - a[:, 1]))),
            "max_abs_dfsimc": float(np.max(np.abs(a[:, 3] - a[:, 1]))),
            "spot_err_median": float(np.median(a[:, 4])), "spot_tol": spot_tol,
            "fallback_rate": float(np.mean([r[5] for r in rows]))}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--patches", type=int, default=20000)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--scale", type=int, default=1,
                    help="make the fields at scale*size and let FSIM downscale them (4: as 1024^2-1080p inputs)")
    ap.add_argument("--seed", type=int, default=40000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--loss-eps", type=float, default=-1.0,
                    help=">= 0: weight each pixel's squared error by max(PC of the pair) + this, "
                         "the weight FSIM gives the pixel; < 0: plain mean squared error")
    ap.add_argument("--lr-halve", type=int, default=3, help="halve the learning rate every this many epochs")
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--out", default="")
    ap.add_argument("--resume", default="")
    ap.add_argument("--eval-pairs", type=int, default=8)
    ap.add_argument("--eval-seed", type=int, default=4100)
    ap.add_argument("--spot-tol", type=float, default=0.02)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--json", default="", help="write the training history and evaluation here")
    args = ap.parse_args(argv)

    ctx = iqm_amd.Context(0)
    sur = iqm_amd.Surrogate(ctx, seed=args.init_seed)
    if args.resume:
        with open(args.resume, "rb") as f:
            sur.load(f.read())
        print(f"[train] resumed at step {sur.state()[0]}", flush=True)
    sur.set_loss_weight(args.loss_eps if args.loss_eps >= 0 else None)
    hist = [] if args.eval_only else train(ctx, sur, args)
    ev = evaluate(ctx, sur, args.eval_pairs, args.eval_seed, args.spot_tol) if args.eval_pairs else {}
    out = {"history": hist, "eval": ev, "step": sur.state()[0]}
    print(json.dumps(out), flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)
    sur.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
