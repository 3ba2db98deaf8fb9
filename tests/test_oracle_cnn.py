"""The CPU restatement of the learned initial guess's network (oracle/cnn_oracle.py),
pinned on its own: the backward pass against central finite differences of the loss in
float64, each adjoint (3 x 3 'valid' im2col, the correlation) against its definition, the
correlation against a direct per-displacement loop, the soft-argmax against a displaced
delta, and the BF16 rounding against known values.  (The network has no counterpart in
the reference.)"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import cnn_oracle as co  # noqa: E402


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3],
                 np.float32)
    r = co.bf16(x)
    assert r[0] == 1.0
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[1] == 1.0
    assert r[2] == 1.0                          # below halfway
    assert r[3] == 1.0 + 2 ** -7                # above halfway
    assert r[4] == -2.5                         # exact
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_conv_cols_adjoint():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 8, 8, 3))
    d = rng.normal(size=(2 * 36, 27))
    lhs = np.sum(co.conv_cols(x) * d)
    rhs = np.sum(x * co.conv_cols_adjoint(d, 2, 8, 3))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_correlation_direct_loop_and_adjoint():
    rng = np.random.default_rng(4)
    nf = rng.normal(size=(2, 16, 16, co.C))
    ng = rng.normal(size=(2, 16, 16, co.C))
    Q = co.correlate(nf, ng)
    # direct: quadrant 3 (qy 1, qx 1) at displacement (dx, dy) = (-2, 4)
    k = (4 + co.R) * (2 * co.R + 1) + (-2 + co.R)
    s = 0.0
    for py in range(8, 11):
        for px in range(8, 11):
            s += np.dot(nf[1, py, px], ng[1, py + 4, px - 2])
    assert abs(Q[1, 3, k] - s / 9.0) < 1e-12
    dQ = rng.normal(size=Q.shape)
    dnf, dng = co.correlate_adjoint(dQ, nf, ng)
    # Q is bilinear: <dQ, Q(nf, ng)> = <dnf, nf> = <dng, ng>
    lhs = np.sum(dQ * Q)
    assert abs(np.sum(dnf * nf) - lhs) < 1e-10 * abs(lhs)
    assert abs(np.sum(dng * ng) - lhs) < 1e-10 * abs(lhs)


def test_softargmax_of_a_peaked_volume():
    V = np.zeros((1, 5, co.ND))
    k = (-3 + co.R) * (2 * co.R + 1) + (2 + co.R)   # (dx, dy) = (2, -3)
    V[:, :, k] = 1.0
    m, w = co.softargmax(V, 200.0)
    assert np.allclose(m[0, :, 0], 8.0) and np.allclose(m[0, :, 1], -12.0)
    m0, _ = co.softargmax(np.zeros((1, 5, co.ND)), 5.0)
    assert np.allclose(m0, 0.0)                     # flat volume: the window's centre


def test_whole_cell_shift_is_found_by_any_encoder():
    """g = f moved by (4a, 4b) pixels: the g features at cell p + (a, b) equal the f
    features at p, and for unit-length features that term is the largest of each
    quadrant's sums (Cauchy-Schwarz), so with a large beta every soft-argmax returns
    (4a, 4b) and the initial head gives u = 4a, v = 4b and zero gradients, whatever
    the (random) encoder."""
    rng = np.random.default_rng(6)
    img = rng.normal(size=(200, 200))
    w = co.init_weights(2).astype(np.float64)
    w[co.nparams() - 67] = 1e4                    # beta
    for a, b in ((0, 0), (3, -2), (-5, 5)):
        f = img[60:60 + co.PATCH, 60:60 + co.PATCH]
        g = img[60 - 4 * b:60 - 4 * b + co.PATCH, 60 - 4 * a:60 - 4 * a + co.PATCH]
        y, c = co.forward(w, np.stack([f, g], -1)[None], dev=False)
        assert np.allclose(c["m"][0, :, 0], 4 * a, atol=1e-6), c["m"]
        assert np.allclose(c["m"][0, :, 1], 4 * b, atol=1e-6), c["m"]
        assert np.allclose(y[0] * co.SCALE, [0, 0, 0, 0, 4 * a, 4 * b], atol=1e-6)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = co.init_weights(3).astype(np.float64)
    w += rng.normal(scale=0.01, size=w.size)      # non-zero biases
    # a smooth pair so that the soft-argmax is neither flat nor saturated
    y0, x0 = np.mgrid[0:co.PATCH, 0:co.PATCH]
    f = np.sin(x0 / 3.1) * np.cos(y0 / 4.3) + 0.3 * rng.normal(size=x0.shape)
    g = np.roll(f, (2, -3), (0, 1))
    patches = np.stack([np.stack([f, g], -1), np.stack([g, f], -1)])
    targets = rng.normal(size=(2, 6))
    w[co.nparams() - 67] = 3.0                    # beta
    loss, grad, _ = co.loss_and_grad(w, patches, targets, dev=False)
    assert grad.shape == w.shape and co.nparams() == w.size
    offs = [int(i) for i in rng.integers(0, 256, 6)]                        # W1
    offs += [256 + int(i) for i in rng.integers(0, 16, 3)]                  # b1
    offs += [272 + int(i) for i in rng.integers(0, 2304, 8)]                # W2
    offs += [2576 + int(i) for i in rng.integers(0, 16, 3)]                 # b2
    offs += [2592]                                                          # beta
    offs += [2593 + int(i) for i in rng.integers(0, 60, 6)]                 # Wl
    offs += [2653 + int(i) for i in range(6)]                               # bl
    assert co.nparams() == 2659
    h = 1e-6
    for i in offs:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (co.loss_and_grad(wp, patches, targets, dev=False)[0]
              - co.loss_and_grad(wm, patches, targets, dev=False)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, grad[i])
