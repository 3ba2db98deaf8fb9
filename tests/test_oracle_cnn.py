"""The CPU restatement of the learned initial guess's network (oracle/cnn_oracle.py),
pinned on its own: the backward pass against central finite differences of the loss in
float64, the im2col adjoint against its definition, and the BF16 rounding against known
values.  (The network has no counterpart in the reference.)"""
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
    d = rng.normal(size=(2 * 16, 27))
    lhs = np.sum(co.conv_cols(x) * d)
    rhs = np.sum(x * co.conv_cols_adjoint(d, 2, 8, 3))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = co.init_weights(3).astype(np.float64)
    w += rng.normal(scale=0.01, size=w.size)      # non-zero biases
    patches = rng.normal(size=(2, 64, 64, 2))
    targets = rng.normal(size=(2, 6))
    loss, g, _ = co.loss_and_grad(w, patches, targets, dev=False)
    assert g.shape == w.shape and co.nparams() == w.size
    # sample parameters from every layer's weights and biases
    offs, o = [], 0
    for n, k in co.SHAPES:
        offs += [o + int(i) for i in rng.integers(0, n * k, 6)]
        offs += [o + n * k + int(i) for i in rng.integers(0, n, 2)]
        o += n * k + n
    h = 1e-6
    for i in offs:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (co.loss_and_grad(wp, patches, targets, dev=False)[0]
              - co.loss_and_grad(wm, patches, targets, dev=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
