"""Trains the learned initial guess (cnn.py) on the GPU and writes its weights.

    python 2d_aldic_amd/cnn_train.py [--out 2d_aldic_amd/data/cnn_weights.npz]
        [--pairs 160] [--size 768] [--step 12] [--steps 12000] [--batch 256] [--seed 0]

Synthetic pairs (cnn.training_pairs), nodes every `step` px at least 48 px inside,
labels from IC-GN started at the known field, kept where IC-GN converged and both
displacement components are within MAX_DISP px (a 64 px patch pair holds no usable
overlap much beyond that); the last four pairs are held out.  Adam
at 1e-3, then a quarter of the steps at 2e-4.  Prints the loss every 500 steps and, on the
held-out pairs, the error of the predicted translation and the IC-GN iterations from the
learned start against the rounded true translation."""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
MAX_DISP = 24.0


def log(msg):
    print(msg, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(ROOT, "2d_aldic_amd", "data", "cnn_weights.npz"))
    ap.add_argument("--pairs", type=int, default=160)
    ap.add_argument("--size", type=int, default=768)
    ap.add_argument("--step", type=int, default=12)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    pkg = importlib.import_module("2d_aldic_amd")
    cnn = importlib.import_module("2d_aldic_amd.cnn")
    t0 = time.perf_counter()
    log(f"[cnn] making {args.pairs} pairs of {args.size}^2")
    pairs = cnn.training_pairs(args.pairs, n=args.size, seed=args.seed)
    log(f"[cnn] {len(pairs)} pairs of {args.size}^2 in {time.perf_counter() - t0:.1f} s")
    ctx = pkg.Context(0)
    assert ctx.cnn_nparams() == sum(n * k + n for n, k in cnn.SHAPES)
    ctx.cnn_set_weights(cnn.init_weights(args.seed))
    ctx.cnn_clear_samples()
    xy = cnn.node_grid(args.size, args.step)
    held = pairs[-4:]
    n_train = 0
    for k, (p, ufun) in enumerate(pairs[:-4]):
        P, ok = cnn.label(ctx, p, ufun, xy)
        ok &= np.abs(P[:, 4:6]).max(axis=1) <= MAX_DISP
        ctx.cnn_add_samples(xy[ok], P[ok])
        n_train += int(ok.sum())
    log(f"[cnn] {n_train} labelled training nodes ({len(xy)} per pair)")
    losses = []
    for phase, (steps, lr) in enumerate([(args.steps, 1e-3), (args.steps // 4, 2e-4)]):
        for c0 in range(0, steps, 500):
            ls = ctx.cnn_train(min(500, steps - c0), args.batch, lr, seed=args.seed * 1000 + phase * 100 + c0)
            losses += ls.tolist()
            log(f"[cnn] phase {phase} step {c0 + len(ls)}: loss {ls[-50:].mean():.5f}")
    w = ctx.cnn_get_weights()
    ctx.cnn_set_weights(w)
    stats = []
    for p, ufun in held:
        P, ok = cnn.label(ctx, p, ufun, xy)
        ok &= np.abs(P[:, 4:6]).max(axis=1) <= MAX_DISP
        Ph = ctx.cnn_predict(xy)
        dU = np.hypot(*(Ph[:, 4:6] - P[:, 4:6])[ok].T)
        dF = np.abs(Ph[:, :4] - P[:, :4])[ok]
        _, _, _, it_c, z_c = ctx.icgn6_start(xy, Ph, 32)
        u, v = ufun(xy[:, 0], xy[:, 1])
        _, _, _, it_r, _ = ctx.icgn6(xy, np.column_stack([np.rint(u), np.rint(v)]).ravel(), 32)
        conv = (it_c >= 1) & (it_c <= cnn.MAX_IT) & (z_c >= cnn.MIN_ZNCC) & ok
        stats.append(dict(rms_u=float(np.sqrt(np.mean(dU ** 2))), within_1px=float(np.mean(dU < 1)),
                          mean_abs_F=float(dF.mean()), converged=float(conv.mean()),
                          iters_learned=float(it_c[conv].mean()), iters_rounded_true=float(it_r[ok].mean())))
        log(f"[cnn] held-out pair: {json.dumps(stats[-1])}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    meta = dict(pairs=args.pairs, size=args.size, step=args.step, steps=args.steps, batch=args.batch,
                seed=args.seed, train_nodes=n_train, final_loss=float(np.mean(losses[-100:])),
                held_out=stats)
    np.savez(args.out, w=w, meta=json.dumps(meta))
    log(f"[cnn] wrote {args.out} in {time.perf_counter() - t0:.1f} s")


if __name__ == "__main__":
    main()
