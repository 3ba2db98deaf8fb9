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
    """(B, 64, 64, 2) patches of the device's normalised images, as the device cuts them."""
    ny, nx = f.shape
    out = np.empty((len(xy), 64, 64, 2), np.float32)
    for b, (x, y) in enumerate(xy):
        x0, y0 = int(np.floor(x)) - 33, int(np.floor(y)) - 33
        xs = np.clip(np.arange(x0, x0 + 64), 0, nx - 1)
        ys = np.clip(np.arange(y0, y0 + 64), 0, ny - 1)
        out[b, :, :, 0] = f[np.ix_(ys, xs)]
        out[b, :, :, 1] = g[np.ix_(ys, xs)]
    return out


def _layer_slices():
    import cnn_oracle as co
    sl, o = [], 0
    for n, k in co.SHAPES:
        sl.append(slice(o, o + n * k))
        sl.append(slice(o + n * k, o + n * k + n))
        o += n * k + n
    return sl


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
    # gives the training outputs
    Pd = gpu_ctx.cnn_predict(xy[rows])
    assert np.allclose(Pd / co.SCALE, ydev, rtol=1e-5, atol=1e-5)


def test_training_lowers_the_loss(gpu_ctx, speckle):
    import cnn_oracle as co
    p = speckle.small_pair(n=256, u0=1.0, v0=0.5, seed=12)
    gpu_ctx.set_images(p.f, p.g)
    gy, gx = np.mgrid[40:216:4, 40:216:4]
    xy = np.column_stack([gx.ravel(), gy.ravel()]).astype(float)
    # a target that depends on the patches: the true translation field of the pair
    P = np.zeros((len(xy), 6))
    P[:, 4], P[:, 5] = 1.0, 0.5
    gpu_ctx.cnn_set_weights(co.init_weights(7))
    gpu_ctx.cnn_clear_samples()
    gpu_ctx.cnn_add_samples(xy, P)
    loss = gpu_ctx.cnn_train(200, 64, 1e-3, seed=3)
    _flush(f"loss {loss[:5].mean():.4f} -> {loss[-5:].mean():.5f}")
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.1 * loss[:5].mean()


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
