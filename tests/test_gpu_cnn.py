"""The learned initial guess on the device (aldic_cnn_*: csrc/cnn.hip) against the CPU
restatement (oracle/cnn_oracle.py, device arithmetic: BF16 GEMM operands, FP32
activations): outputs, loss and every layer's gradient on patch pairs cut from a speckle
pair, then a short training run that must lower the loss, and IC-GN started from a full
affine warp (aldic_icgn6_start)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))


def _flush(msg):
    print(msg, flush=True)


def _patches(f, g, xy):
    """(B, 72, 72, 2) inputs of the device's normalised images, as the device cuts them:
    the 64 x 64 patch from pixel x0 - 32 (1-based node x0) and a 4-pixel ring."""
    ny, nx = f.shape
    out = np.empty((len(xy), 72, 72, 2), np.float32)
    for b, (x, y) in enumerate(xy):
        x0, y0 = int(np.floor(x)) - 37, int(np.floor(y)) - 37
        xs = np.clip(np.arange(x0, x0 + 72), 0, nx - 1)
        ys = np.clip(np.arange(y0, y0 + 72), 0, ny - 1)
        out[b, :, :, 0] = f[np.ix_(ys, xs)]
        out[b, :, :, 1] = g[np.ix_(ys, xs)]
    return out


def _layer_slices():
    """W1 b1 W2 b2 beta Wl bl in the flat parameter vector"""
    sizes = [256, 16, 2304, 16, 1, 60, 6]
    o = np.concatenate([[0], np.cumsum(sizes)])
    return [slice(o[k], o[k + 1]) for k in range(len(sizes))]


def test_forward_and_gradients_match_oracle(gpu_ctx, speckle):
    import cnn_oracle as co
    p = speckle.small_pair(n=200, u0=2.3, v0=-1.4, ux=0.01, vy=-0.008, seed=11)
    gpu_ctx.set_images(p.f, p.g)
    f, g, _, _ = gpu_ctx.get_images()
    rng = np.random.default_rng(2)
    xy = np.column_stack([rng.integers(40, 160, 12), rng.integers(40, 160, 12)]).astype(float)
    xy[0] = (20.0, 30.0)                       # clamped at the border
    P = rng.normal(size=(12, 6)) * np.array([0.02, 0.02, 0.02, 0.02, 4.0, 4.0])
    w0 = co.init_weights(5)
    w0[co.nparams() - 6:] = 0.1                # non-zero output biases
    w0[co.nparams() - 67] = 4.0                # beta: soft-argmax neither flat nor saturated
    assert gpu_ctx.cnn_nparams() == co.nparams() == w0.size
    gpu_ctx.cnn_set_weights(w0)
    gpu_ctx.cnn_clear_samples()
    gpu_ctx.cnn_add_samples(xy, P)
    rows = np.array([3, 0, 7, 11, 5], np.int32)
    gdev, ldev, ydev = gpu_ctx.cnn_grad(rows)
    tgt = (P / co.SCALE).astype(np.float32).astype(np.float64)[rows]
    patches = _patches(f, g, xy[rows])
    lo, go, yo = co.loss_and_grad(w0, patches, tgt, dev=True)
    ey = np.linalg.norm(ydev - yo) / np.linalg.norm(yo)
    _flush(f"outputs rel {ey:.2e}, loss {ldev:.6f} / {lo:.6f}")
    assert ey < 2e-3
    assert abs(ldev - lo) <= 2e-3 * lo
    for k, s in enumerate(_layer_slices()):
        e = np.linalg.norm(gdev[s] - go[s]) / max(np.linalg.norm(go[s]), 1e-30)
        _flush(f"grad block {k}: rel {e:.2e}")
        assert e < 2e-2, (k, e)
    # the patch reads from the images equal the stored ones: predicting on the same nodes
    # (per-patch path: their starts do not share one 4-pixel lattice) gives the
    # training outputs
    Pd = gpu_ctx.cnn_predict(xy[rows])
    assert np.allclose(Pd / co.SCALE, ydev, rtol=1e-5, atol=1e-5)


def test_lattice_maps_equal_per_patch_features(gpu_ctx, speckle):
    """On a grid of step 8 the nodes' inputs share one 4-pixel lattice and the encoder
    runs once per image; adding one node off the lattice sends every node through the
    per-patch path.  Both give the same outputs (same cells, same arithmetic)."""
    import cnn_oracle as co
    p = speckle.small_pair(n=240, u0=3.2, v0=-2.6, ux=0.01, vy=-0.01, seed=14)
    gpu_ctx.set_images(p.f, p.g)
    gy, gx = np.mgrid[30:211:8, 25:216:8]      # nodes near the border too (clamped reads)
    xy = np.column_stack([gx.ravel(), gy.ravel()]).astype(float)
    w0 = co.init_weights(9)
    w0[co.nparams() - 67] = 6.0
    gpu_ctx.cnn_set_weights(w0)
    Pl = gpu_ctx.cnn_predict(xy)
    Pp = gpu_ctx.cnn_predict(np.vstack([xy, [[101.0, 99.0]]]))[:-1]
    err = np.max(np.abs(Pl - Pp) / co.SCALE)
    _flush(f"lattice vs per-patch: max scaled difference {err:.2e} over {len(xy)} nodes")
    assert err < 1e-4
    # and against the oracle's forward pass on a few nodes
    f, g, _, _ = gpu_ctx.get_images()
    sel = np.array([0, 7, 100, len(xy) - 1])
    yo, _ = co.forward(w0, _patches(f, g, xy[sel]), dev=True)
    assert np.allclose(Pl[sel] / co.SCALE, yo, rtol=2e-3, atol=2e-3)


def test_training_lowers_the_loss(gpu_ctx, speckle):
    """Targets from the pair's true field (translation 5.3, -3.7 px plus gradients): the
    loss of the randomly initialised encoder falls to under half in 300 Adam steps."""
    import cnn_oracle as co
    G = np.array([0.012, -0.008, 0.006, 0.015])          # u_x, v_x, u_y, v_y
    p = speckle.small_pair(n=256, u0=5.3, v0=-3.7, ux=G[0], vx=G[1], uy=G[2], vy=G[3], seed=12)
    gpu_ctx.set_images(p.f, p.g)
    gy, gx = np.mgrid[60:200:4, 60:200:4]
    xy = np.column_stack([gx.ravel(), gy.ravel()]).astype(float)
    c = (256 + 1) / 2.0
    P = np.zeros((len(xy), 6))
    P[:, :4] = G
    P[:, 4] = 5.3 + G[0] * (xy[:, 0] - c) + G[2] * (xy[:, 1] - c)
    P[:, 5] = -3.7 + G[1] * (xy[:, 0] - c) + G[3] * (xy[:, 1] - c)
    gpu_ctx.cnn_set_weights(co.init_weights(7))
    gpu_ctx.cnn_clear_samples()
    gpu_ctx.cnn_add_samples(xy, P)
    loss = gpu_ctx.cnn_train(300, 64, 1e-3, seed=3)
    _flush(f"loss {loss[:5].mean():.4f} -> {loss[-5:].mean():.5f}")
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.5 * loss[:5].mean()


def test_icgn6_start_from_affine_warp(gpu_ctx, speckle):
    """Started from the pair's affine warp, IC-GN converges to the answer it reaches from
    zero gradients; started from u_x = 8 its first warp leaves the image (status 101,
    funICGN.m:118-121), which only happens if the start is used."""
    p = speckle.small_pair(n=160, u0=1.6, v0=-0.9, ux=0.02, uy=-0.015, vx=0.01, vy=0.025,
                           seed=13)
    gpu_ctx.set_images(p.f, p.g)
    c = (160 + 1) / 2.0                        # the warp's centre (small_pair)
    xy = np.array([[80.0, 80.0], [60.0, 90.0], [95.0, 68.0]])   # whole-pixel nodes
    P0 = np.zeros((3, 6))
    P0[:, 0], P0[:, 1], P0[:, 2], P0[:, 3] = 0.02, 0.01, -0.015, 0.025
    P0[:, 4] = 1.6 + 0.02 * (xy[:, 0] - c) - 0.015 * (xy[:, 1] - c)
    P0[:, 5] = -0.9 + 0.01 * (xy[:, 0] - c) + 0.025 * (xy[:, 1] - c)
    U, F, H, it, z = gpu_ctx.icgn6_start(xy, P0, 20, tol=1e-2)
    Uz, Fz, Hz, itz, zz = gpu_ctx.icgn6(xy, np.rint(P0[:, 4:6]).ravel(), 20, tol=1e-2)
    Pb = P0.copy()
    Pb[:, 0] = 8.0
    Ub, Fb, Hb, itb, zb = gpu_ctx.icgn6_start(xy, Pb, 20, tol=1e-2)
    _flush(f"iterations from the warp {it}, from zero gradients {itz}, from u_x 8 {itb}")
    assert np.all((it >= 1) & (it <= 100)) and np.all((itz >= 1) & (itz <= 100))
    assert np.all(itb == 101)
    assert np.allclose(U, Uz, atol=2e-2) and np.allclose(F, Fz, atol=2e-3)
    assert np.allclose(U.reshape(-1, 2), P0[:, 4:6], atol=0.05)


def test_local_icgn_learned_start_with_shipped_weights(gpu_ctx, speckle):
    """aldic_local_icgn on a 12-bit 1024^2 pair (tension plus a shear band, like the
    4096^2 workload): from the search it equals aldic_integer_search + aldic_icgn6; from
    the shipped network's warp every converged subset lands on the same displacement
    (IC-GN's optimum, compared at the tighter stop tolerance 1e-3 so that two converged
    runs agree to about that) in at most 0.75 more iterations on average (the warp is
    within about 1 px, the search's peak within 0.5 px); with an unreachable ZNCC bar every
    subset falls back and the result is the search start's (FP32 search peaks on the
    fallback)."""
    import importlib
    cnn = importlib.import_module("2d_aldic_amd.cnn")
    p = speckle.config3(seed=3, n=1024, n_dots=37_500)
    gpu_ctx.set_images(p.f, p.g)
    cnn.load_weights(gpu_ctx)
    grid, ws, step, s = [1, 1024, 1, 1024], 32, 8, [24, 24]
    r0 = gpu_ctx.local_icgn(grid, ws, step, s, start="search", tol=1e-3)
    sr = gpu_ctx.integer_search(grid, ws, step, s)
    xy = np.column_stack([sr["x"].ravel(), sr["y"].ravel()])
    U0 = np.column_stack([sr["u"].ravel(), sr["v"].ravel()]).ravel()
    U, F, H, it, z = gpu_ctx.icgn6(xy, U0, ws, tol=1e-3)
    assert np.array_equal(np.column_stack([r0["x"], r0["y"]]), xy)
    assert np.array_equal(r0["iters"], it) and np.array_equal(r0["U"], U)
    r1 = gpu_ctx.local_icgn(grid, ws, step, s, start="learned", tol=1e-3)
    ok0 = (r0["iters"] >= 1) & (r0["iters"] <= 100)
    ok1 = (r1["iters"] >= 1) & (r1["iters"] <= 100)
    both = ok0 & ok1
    dU = np.abs(r1["U"] - r0["U"]).reshape(-1, 2)[both]
    _flush(f"S {len(ok0)}: converged search {ok0.mean():.4f}, learned {ok1.mean():.4f}, "
           f"fallback {r1['n_fallback']}, iters {r0['iters'][ok0].mean():.2f} / "
           f"{r1['iters'][ok1].mean():.2f}, max |dU| {dU.max():.2e}, "
           f"times search {r0['times_ms'][:2]} learned {r1['times_ms']}")
    assert ok1.mean() >= ok0.mean() - 1e-3
    assert r1["n_fallback"] == int(r1["fallback"].sum()) <= 0.02 * len(ok0)
    assert dU.max() < 5e-3
    assert r1["iters"][ok1].mean() <= r0["iters"][ok0].mean() + 0.75
    rf = gpu_ctx.local_icgn(grid, ws, step, s, start="learned", tol=1e-3, min_zncc=2.0)
    assert rf["n_fallback"] == len(ok0) and rf["fallback"].all()
    same = (rf["iters"] == r0["iters"])
    assert same.mean() > 0.99
    assert np.max(np.abs(rf["U"] - r0["U"]).reshape(-1, 2)[same & ok0]) < 1e-3


def test_frame_with_learned_start(gpu_ctx, speckle):
    """The device frame (aldic_frame_run) with learned_start: LocalICGN starts from the
    shipped network's warp instead of the integer search (FP32 search and IC-GN again on
    the subsets that miss status 1..20 or ZNCC 0.9), then the same ADMM steps.  On the
    12-bit 1024^2 pair it converges the same subsets as the searched frame, LocalICGN's U
    agrees to IC-GN's stop tolerance (1e-3 here) and the final U and F to about that."""
    import importlib
    cnn = importlib.import_module("2d_aldic_amd.cnn")
    p = speckle.config3(seed=3, n=1024, n_dots=37_500)
    gpu_ctx.set_images(p.f, p.g)
    cnn.load_weights(gpu_ctx)
    gpu_ctx.frame_stage(p.f, p.g)
    grid, ws, step, s = [1, 1024, 1, 1024], 32, 8, [24, 24]
    res = {}
    for learned in (False, True):
        _flush(f"frame learned_start={learned}")
        gpu_ctx.frame_setup(ws, step, s, grid, tol=1e-3, learned_start=learned)
        gpu_ctx.frame_run()
        res[learned] = gpu_ctx.frame_get()
        res[learned]["fallback"] = gpu_ctx.frame_counters()["learned_fallback"]
    r0, r1 = res[False], res[True]
    assert r0["fallback"] == -1 and 0 <= r1["fallback"] <= 0.02 * r0["iters6"].size
    ok0 = (r0["iters6"] >= 1) & (r0["iters6"] <= 100)
    ok1 = (r1["iters6"] >= 1) & (r1["iters6"] <= 100)
    both = ok0 & ok1
    d6 = np.abs(r1["U6"] - r0["U6"]).reshape(-1, 2)[both]
    dU = np.abs(r1["U"] - r0["U"]).reshape(-1, 2)
    dF = np.abs(r1["F"] - r0["F"])
    _flush(f"converged {ok0.mean():.4f} / {ok1.mean():.4f}, fallback {r1['fallback']}, "
           f"max |dU6| {d6.max():.2e}, max |dU| {dU.max():.2e}, rms {np.sqrt(np.mean(dU ** 2)):.2e}, "
           f"max |dF| {dF.max():.2e}")
    assert ok1.mean() >= ok0.mean() - 1e-3
    assert d6.max() < 5e-3
    assert dU.max() < 2e-2 and np.sqrt(np.mean(dU ** 2)) < 2e-3
    assert dF.max() < 1e-3
