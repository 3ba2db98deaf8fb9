"""Trains the learned initial guess (cnn.py) on the GPU and writes its weights.

    python 2d_aldic_amd/cnn_train.py [--out 2d_aldic_amd/data/cnn_weights.npz]
        [--pairs 16000] [--size 192] [--step 12] [--steps 20000] [--batch 256] [--seed 0]

Synthetic pairs (cnn.training_pairs), each with its own random field, nodes every `step`
px at least 48 px inside, labels from IC-GN started at the known field, kept where IC-GN
converged and both displacement components are within MAX_DISP px (the reach of the
cost volumes); the last 2 % of the pairs are held out.  Adam at 1e-3, then a quarter of
the steps at 2e-4.  Prints the loss every 500 steps and, on the
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


def log(msg):
    print(msg, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(ROOT, "2d_aldic_amd", "data", "cnn_weights.npz"))
    ap.add_argument("--pairs", type=int, default=16000)
    ap.add_argument("--size", type=int, default=192)
    ap.add_argument("--step", type=int, default=12)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    pkg = importlib.import_module("2d_aldic_amd")
    cnn = importlib.import_module("2d_aldic_amd.cnn")
    t0 = time.perf_counter()
    log(f"[cnn] making {args.pairs} pairs of {args.size}^2")
    pairs = []
    for c0 in range(0, args.pairs, 500):
        pairs += cnn.training_pairs(min(500, args.pairs - c0), n=args.size, seed=args.seed,
                                    first=c0, workers=16)
        log(f"[cnn] {len(pairs)} pairs of {args.size}^2 in {time.perf_counter() - t0:.1f} s")
    ctx = pkg.Context(0)
    assert ctx.cnn_nparams() == cnn.NPARAMS
    ctx.cnn_set_weights(cnn.init_weights(args.seed))
    ctx.cnn_clear_samples()
    xy = cnn.node_grid(args.size, args.step)
    n_held = max(4, args.pairs // 50)
    held = pairs[-n_held:]
    n_train = 0
    for k, (p, ufun) in enumerate(pairs[:-n_held]):
        P, ok = cnn.label(ctx, p, ufun, xy)
        ok &= np.abs(P[:, 4:6]).max(axis=1) <= cnn.MAX_DISP
        ctx.cnn_add_samples(xy[ok], P[ok])
        n_train += int(ok.sum())
        if (k + 1) % 1000 == 0:
            log(f"[cnn] labelled {k + 1} pairs")
    log(f"[cnn] {n_train} labelled training nodes ({len(xy)} per pair)")
    losses = []
    for phase, (steps, lr) in enumerate([(args.steps, 1e-3), (args.steps // 4, 2e-4)]):
        for c0 in range(0, steps, 500):
            ls = ctx.cnn_train(min(500, steps - c0), args.batch, lr, seed=args.seed * 1000 + phase * 100 + c0)
            losses += ls.tolist()
            log(f"[cnn] phase {phase} step {c0 + len(ls)}: loss {ls[-50:].mean():.5f}")
    w = ctx.cnn_get_weights()
    ctx.cnn_set_weights(w)
    dU, dF, it_c, it_r, conv, n_ok = [], [], [], [], [], 0
    for p, ufun in held:
        P, ok = cnn.label(ctx, p, ufun, xy)
        ok &= np.abs(P[:, 4:6]).max(axis=1) <= cnn.MAX_DISP
        Ph = ctx.cnn_predict(xy)
        dU.append(np.hypot(*(Ph[:, 4:6] - P[:, 4:6])[ok].T))
        dF.append(np.abs(Ph[:, :4] - P[:, :4])[ok].ravel())
        _, _, _, itc, zc = ctx.icgn6_start(xy, Ph, 32)
        u, v = ufun(xy[:, 0], xy[:, 1])
        _, _, _, itr, _ = ctx.icgn6(xy, np.column_stack([np.rint(u), np.rint(v)]).ravel(), 32)
        cv = (itc >= 1) & (itc <= cnn.MAX_IT) & (zc >= cnn.MIN_ZNCC)
        conv.append(cv[ok])
        it_c.append(itc[ok & cv])
        it_r.append(itr[ok])
    dU, dF = np.concatenate(dU), np.concatenate(dF)
    conv = np.concatenate(conv)
    stats = dict(nodes=int(dU.size), rms_u=float(np.sqrt(np.mean(dU ** 2))),
                 within_1px=float(np.mean(dU < 1)), within_2px=float(np.mean(dU < 2)),
                 mean_abs_F=float(dF.mean()), converged=float(conv.mean()),
                 iters_learned=float(np.concatenate(it_c).mean()),
                 iters_rounded_true=float(np.concatenate(it_r).mean()))
    log(f"[cnn] held-out pairs ({len(held)}): {json.dumps(stats)}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    meta = dict(pairs=args.pairs, size=args.size, step=args.step, steps=args.steps, batch=args.batch,
                seed=args.seed, train_nodes=n_train, final_loss=float(np.mean(losses[-100:])),
                held_out=stats)
    np.savez(args.out, w=w, meta=json.dumps(meta))
    log(f"[cnn] wrote {args.out} in {time.perf_counter() - t0:.1f} s")


if __name__ == "__main__":
    main()
