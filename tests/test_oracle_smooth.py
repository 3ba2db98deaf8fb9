"""Pins oracle/smooth_oracle.py (regularizeNd.m, funSmoothStrainQuadtree.m) with what the
mathematics fixes: bilinear data are fitted exactly (they are the second differences'
null space and the interpolation reproduces them); the result does not change when the
coordinates and the grid are scaled or shifted (the regularizer's range^2 scaling);
a very stiff fit tends to the least-squares bilinear fit of the data; a very loose one
returns data given at every grid node."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))

import smooth_oracle as so  # noqa: E402


def _bilin(x, y):
    return 0.3 - 0.02 * x + 0.05 * y + 0.001 * x * y


def _pts(rng, n, xg, yg, on_grid=0.5):
    x = rng.uniform(xg[0], xg[-1], n)
    y = rng.uniform(yg[0], yg[-1], n)
    k = rng.random(n) < on_grid
    x[k] = rng.choice(xg, k.sum())
    y[k] = rng.choice(yg, k.sum())
    return x, y


def test_bilinear_data_fitted_exactly():
    rng = np.random.default_rng(0)
    xg = np.linspace(1.0, 41.0, 11)
    yg = np.linspace(-3.0, 25.0, 8)
    x, y = _pts(rng, 60, xg, yg)
    Z = so.regularize_nd(x, y, _bilin(x, y), xg, yg, 1e-2)
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    assert Z.shape == (xg.size, yg.size)
    assert np.max(np.abs(Z - _bilin(X, Y))) < 1e-9


def test_invariant_under_scaling_and_shift():
    rng = np.random.default_rng(1)
    xg = np.arange(0.0, 90.0, 8.0)
    yg = np.arange(0.0, 70.0, 8.0)
    x, y = _pts(rng, 50, xg, yg)
    z = np.sin(x / 17.0) + np.cos(y / 11.0)
    Z1 = so.regularize_nd(x, y, z, xg, yg, 1e-3)
    Z2 = so.regularize_nd(2.5 * x + 7.0, 0.5 * y - 3.0, z, 2.5 * xg + 7.0, 0.5 * yg - 3.0, 1e-3)
    assert np.max(np.abs(Z1 - Z2)) < 1e-9


def test_stiff_fit_tends_to_bilinear_least_squares():
    rng = np.random.default_rng(2)
    xg = np.linspace(0.0, 30.0, 7)
    yg = np.linspace(0.0, 20.0, 6)
    x, y = _pts(rng, 40, xg, yg, on_grid=0.0)
    z = rng.normal(size=40)
    Z = so.regularize_nd(x, y, z, xg, yg, 1e3)   # stiffer: the normal equations lose digits
    M = np.column_stack([np.ones_like(x), x, y, x * y])
    c = np.linalg.lstsq(M, z, rcond=None)[0]
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    ref = c[0] + c[1] * X + c[2] * Y + c[3] * X * Y
    assert np.max(np.abs(Z - ref)) < 1e-5 * (1 + np.abs(ref).max())


def test_loose_fit_returns_grid_data():
    rng = np.random.default_rng(3)
    xg = np.arange(0.0, 50.0, 5.0)
    yg = np.arange(0.0, 35.0, 5.0)
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    z = rng.normal(size=X.size)
    Z = so.regularize_nd(X.ravel(), Y.ravel(), z, xg, yg, 1e-9)
    assert np.max(np.abs(Z.ravel() - z)) < 1e-6


def test_smooth_strain_quadtree_keeps_bilinear_interior_one_pass_and_rejects_off_grid():
    xs = np.arange(0.0, 81.0, 8.0)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel()])
    F = np.zeros(4 * len(coords))
    for c in range(4):
        F[c::4] = (c + 1) * 0.01 + 1e-4 * coords[:, 0] - 2e-4 * coords[:, 1]
    out = so.smooth_strain_quadtree(F, coords, 8.0)
    # linear fields survive regularizeNd exactly and the symmetric blur in the interior;
    # the edge rows change (the blur renormalises over the in-image part only) and the
    # second pass carries that inwards only slightly
    inner = (coords[:, 0] > 16) & (coords[:, 0] < 64) & (coords[:, 1] > 16) & (coords[:, 1] < 64)
    for c in range(4):
        assert np.max(np.abs(out[c::4][inner] - F[c::4][inner])) < 2e-5
    bad = coords.copy()
    bad[5, 0] += 1.0
    try:
        so.smooth_strain_quadtree(F, bad, 8.0)
    except ValueError:
        pass
    else:
        raise AssertionError("off-grid node accepted")
