"""Learned initial guess for the 6-DOF IC-GN (a start in place of the integer search).

A siamese correlation network on the BF16 matrix cores (csrc/cnn.hip; C ABI aldic_cnn_*)
maps each subset's 64 x 64 reference and deformed patches of the normalised images,
centred on the node (plus a 4-pixel context ring), to the six IC-GN parameters (u_x, v_x,
u_y, v_y, u, v): a shared two-layer encoder on 4 x 4 pixel cells, normalised features,
cost volumes over displacements of up to 5 cells (20 px) for the four quadrants of the
central 24 x 24 pixels, a soft-argmax of each and a dense 10 -> 6 head.  IC-GN then
starts from that warp (aldic_icgn6_start); the plain integer search is the fallback for
every subset whose IC-GN did not converge within 20 iterations or ends with ZNCC below
0.9, and converged IC-GN decides every displacement.

Training data come from the synthetic speckle warper (speckle.make_pair): many small
image pairs, each with its own random translation, displacement gradient and smooth
sinusoidal perturbation; each node is labelled by IC-GN started from the known field
rounded to whole pixels (converged subsets only).  The trained weights ship as
data/cnn_weights.npz, written by cnn_train.py.  No counterpart in the reference (its
initial guess is the FFT search of IntegerSearch.m).
"""
from __future__ import annotations

import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "cnn_weights.npz")
MARGIN = 48          # nodes at least this far from the border (patch 32 + ring 4 + window 12)
MAX_IT = 20          # fallback when IC-GN from the learned start needs more
MIN_ZNCC = 0.9       # or ends below this ZNCC
MAX_DISP = 19.0      # labels kept within the cost volumes' reach (5 cells of 4 px)
CH = 16              # encoder channels
NPARAMS = CH * 16 + CH + CH * 9 * CH + CH + 1 + 6 * 10 + 6
SCALE = np.array([0.05, 0.05, 0.05, 0.05, 4.0, 4.0])


def init_weights(seed=0):
    """Flat FP32 parameters in the device's order W1 b1 W2 b2 beta Wl bl: He-normal
    encoder, zero biases, beta = 20, and a head that starts as the geometric estimate
    (global soft-argmax -> u, v; quadrant differences over their 12 px spacing ->
    gradients)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(scale=np.sqrt(2.0 / 16), size=(CH, 16))
    W2 = rng.normal(scale=np.sqrt(1.0 / (9 * CH)), size=(CH, 9 * CH))
    Wl = np.zeros((6, 10))
    Wl[4, 0] = 1.0 / SCALE[4]
    Wl[5, 1] = 1.0 / SCALE[5]
    for comp, row_x, row_y in ((0, 0, 2), (1, 1, 3)):
        for q in range(4):
            qy, qx = divmod(q, 2)
            Wl[row_x, 2 + 2 * q + comp] = (1.0 if qx else -1.0) / 24.0 / SCALE[row_x]
            Wl[row_y, 2 + 2 * q + comp] = (1.0 if qy else -1.0) / 24.0 / SCALE[row_y]
    w = np.concatenate([W1.ravel(), np.zeros(CH), W2.ravel(), np.zeros(CH), [20.0],
                        Wl.ravel(), np.zeros(6)])
    return w.astype(np.float32)


def load_weights(ctx, path=WEIGHTS):
    """Sets the shipped weights on ctx (raises when the file is missing)."""
    with np.load(path, allow_pickle=False) as d:
        w = np.asarray(d["w"], np.float32)
    ctx.cnn_set_weights(w)
    return w


def random_field(rng, n, trans=16.0, grad=0.05, wiggle=3.0):
    """u(X, Y), v(X, Y): translation U[-trans, trans], gradient U[-grad, grad] about the
    image centre, plus a sinusoid of amplitude U[0, wiggle] px and wavelength U[150, 400]
    px along a random direction (per component)."""
    c = (n + 1) / 2.0
    t = rng.uniform(-trans, trans, 2)
    G = rng.uniform(-grad, grad, (2, 2))
    amp = rng.uniform(0, wiggle, 2)
    lam = rng.uniform(150, 400, 2)
    ang = rng.uniform(0, np.pi, 2)
    ph = rng.uniform(0, 2 * np.pi, 2)

    def ufun(X, Y):
        dx, dy = X - c, Y - c
        out = []
        for k in range(2):
            s = np.cos(ang[k]) * dx + np.sin(ang[k]) * dy
            out.append(t[k] + G[k, 0] * dx + G[k, 1] * dy
                       + amp[k] * np.sin(2 * np.pi * s / lam[k] + ph[k]))
        return out[0], out[1]

    return ufun


def _make(args):
    import importlib
    speckle = importlib.import_module("2d_aldic_amd.speckle")
    n, seed = args
    rng = np.random.default_rng(seed)
    ufun = random_field(rng, n)
    sigma = rng.uniform(1.5, 3.0)
    bits = 8 if rng.random() < 0.5 else 12
    noise = rng.uniform(0.0, 3.0) * (1 if bits == 8 else 16)
    dens = 0.045 * (2.0 / sigma) ** 2
    p = speckle.make_pair(n, n, int(dens * n * n), sigma, ufun, noise, bits, seed=seed)
    return p.f, p.g, p.params


def training_pairs(n_pairs, n=192, seed=0, first=0, workers=8):
    """Pairs first .. first + n_pairs - 1 of a seeded sequence of synthetic pairs (n x n)
    with their fields [(pair, ufun)], made in parallel processes.  The field of each pair
    is the first draw of its own generator, so the parent rebuilds it from the seed
    (closures do not cross process boundaries)."""
    args = [(n, seed * 100003 + k) for k in range(first, first + n_pairs)]
    if workers > 1:
        from multiprocessing import get_context
        with get_context("spawn").Pool(workers) as pool:
            pairs = pool.map(_make, args)
    else:
        pairs = [_make(a) for a in args]
    import importlib
    speckle = importlib.import_module("2d_aldic_amd.speckle")
    out = []
    for (f, g, params), (_, s) in zip(pairs, args):
        ufun = random_field(np.random.default_rng(s), n)
        out.append((speckle.SpecklePair(f, g, ufun, params), ufun))
    return out


def node_grid(n, step, margin=MARGIN):
    xs = np.arange(margin, n - margin + 1, step, dtype=float)
    X, Y = np.meshgrid(xs, xs)
    return np.column_stack([X.ravel(), Y.ravel()])


def label(ctx, p, ufun, xy, ws=32, tol=1e-2):
    """IC-GN from the known field rounded to whole pixels: labels P (u_x, v_x, u_y, v_y,
    u, v) of the converged nodes with ZNCC >= MIN_ZNCC, and the mask of those nodes."""
    ctx.set_images(p.f, p.g)
    u, v = ufun(xy[:, 0], xy[:, 1])
    U0 = np.column_stack([np.rint(u), np.rint(v)]).ravel()
    U, F, _, it, z = ctx.icgn6(xy, U0, ws, tol)
    ok = (it >= 1) & (it <= 100) & (z >= MIN_ZNCC)
    P = np.column_stack([F.reshape(-1, 4), U.reshape(-1, 2)])
    return P, ok


def local_icgn(ctx, xy, ws, tol=1e-2, grid=None, s=None, step=None):
    """LocalICGN with the learned start.  Returns (U, F, iters, zncc, info): IC-GN from the
    network's warp; the subsets that need more than MAX_IT iterations, fail, or end below
    MIN_ZNCC are redone from the plain integer search (aldic_integer_search over `grid`
    with search radius `s` and `step`, whose nodes must be xy in the search's order).
    info holds the stage times (ms) and the fallback count."""
    t0 = time.perf_counter()
    P = ctx.cnn_predict(xy)
    t1 = time.perf_counter()
    U, F, H, it, z = ctx.icgn6_start(xy, P, ws, tol)
    t2 = time.perf_counter()
    bad = ~((it >= 1) & (it <= MAX_IT) & (z >= MIN_ZNCC))
    info = dict(cnn_ms=1e3 * (t1 - t0), icgn_ms=1e3 * (t2 - t1), fallback=int(bad.sum()),
                search_ms=0.0, icgn_fallback_ms=0.0)
    if bad.any():
        if grid is None:
            raise ValueError("local_icgn: subsets need the fallback search but no grid is given")
        t3 = time.perf_counter()
        r = ctx.integer_search(grid, ws, step, s)
        if r["x"].size != len(xy) or not np.allclose(
                np.column_stack([r["x"].ravel(), r["y"].ravel()]), xy):
            raise ValueError("local_icgn: the search grid's nodes are not xy")
        U0 = np.column_stack([r["u"].ravel(), r["v"].ravel()])[bad].ravel()
        t4 = time.perf_counter()
        Ub, Fb, Hb, itb, zb = ctx.icgn6(xy[bad], U0, ws, tol)
        t5 = time.perf_counter()
        U = U.reshape(-1, 2)
        F = F.reshape(-1, 4)
        U[bad] = Ub.reshape(-1, 2)
        F[bad] = Fb.reshape(-1, 4)
        it = it.copy()
        z = z.copy()
        it[bad] = itb
        z[bad] = zb
        U, F = U.ravel(), F.ravel()
        info.update(search_ms=1e3 * (t4 - t3), icgn_fallback_ms=1e3 * (t5 - t4))
    return U, F, it, z, info


__all__ = ["NPARAMS", "init_weights", "load_weights", "random_field", "training_pairs",
           "node_grid", "label", "local_icgn", "WEIGHTS", "MAX_IT", "MIN_ZNCC"]
