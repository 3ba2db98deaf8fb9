"""The multigrid integer search on the device (aldic_integer_search_mg: csrc/search_mg.hip,
funIntegerSearchMg.m:3-285) against the CPU restatement (oracle/mg_oracle.py) on the
device's own normalised images: level count, every final region's centre and peak (NaN
where the reference's catch fires), and the springs fit onto the grid."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))


def _flush(msg):
    print(msg, flush=True)


def _compare(d, r):
    """device result d against oracle result r"""
    assert d["levels"] == r["levels"]
    assert np.array_equal(d["tempx"], r["tempx"]) and np.array_equal(d["tempy"], r["tempy"])
    nan_d, nan_r = np.isnan(d["tempu"]), np.isnan(r["tempu"])
    assert np.array_equal(nan_d, nan_r) and np.array_equal(np.isnan(d["tempv"]), nan_r)
    ok = ~nan_r
    # FP64 sums in another order: equal to ~1e-12 except where findpeak's 1e-3 rounding
    # of the paraboloid offset falls the other way (one step of 1e-3, rare)
    for k in ("tempu", "tempv"):
        dd = np.abs(d[k][ok] - r[k][ok])
        assert dd.max() <= 1.001e-3 + 1e-9, (k, dd.max())
        assert np.mean(dd > 1e-9) <= 0.02, (k, np.mean(dd > 1e-9))
    dp = np.abs(d["tempphi"] - r["tempphi"])
    assert dp.max() <= 1e-4 and np.median(dp) <= 1e-12
    assert np.array_equal(d["x"][0], r["xlist"]) and np.array_equal(d["y"][:, 0], r["ylist"])
    # the grid fit: CG to a relative residual of 1e-12 against a direct solve
    for k in ("u", "v", "ccmax"):
        assert d[k].shape == r[k].shape
        dd = np.abs(d[k] - r[k])
        assert dd.max() <= 2e-3, (k, dd.max())
        assert np.median(dd) <= 1e-7, (k, np.median(dd))


def test_search_mg_matches_oracle_square(gpu_ctx, speckle):
    import mg_oracle as mg
    _flush("search_mg: 200 x 200 affine pair, ws 32, step 16")
    p = speckle.small_pair(n=200, u0=2.3, v0=-1.4, ux=0.01, vy=-0.008, seed=11)
    gpu_ctx.set_images(p.f, p.g)
    f, g, _, _ = gpu_ctx.get_images()
    d = gpu_ctx.integer_search_mg((1, 200, 1, 200), 32, 16)
    r = mg.search_mg(f, g, [1, 200], [1, 200], 32, 16)
    _compare(d, r)
    assert d["n_nan"] == 0 and d["levels"] == 4
    # the displacement at the grid nodes is the affine field to the search's accuracy
    c = 100.5
    assert np.max(np.abs(d["u"] - (2.3 + 0.01 * (d["x"] - c)))) < 0.5
    assert np.max(np.abs(d["v"] - (-1.4 - 0.008 * (d["y"] - c)))) < 0.5


def test_search_mg_matches_oracle_rectangular_large_shift(gpu_ctx, speckle):
    """A rectangular ROI (FFT lengths differ per axis) and a shift that moves regions past
    the border gap (clamps of :59-66 / :202-229, odd-width fixes)."""
    import mg_oracle as mg
    _flush("search_mg: 240 x 176 translation (6.6, -5.2), ws 24, step 12")
    p = speckle.make_pair(240, 176, int(0.045 * 240 * 176), 2.0, lambda X, Y: (6.6 + 0 * X, -5.2 + 0 * Y),
                          0.0, 8, seed=5)
    gpu_ctx.set_images(p.f, p.g)
    f, g, _, _ = gpu_ctx.get_images()
    d = gpu_ctx.integer_search_mg((5, 236, 3, 170), 24, 12)
    r = mg.search_mg(f, g, [5, 236], [3, 170], 24, 12)
    _compare(d, r)
    ok = ~np.isnan(d["tempu"])
    assert np.mean(np.abs(d["tempu"][ok] - 6.6) < 0.5) > 0.9


def test_search_mg_flat_block_gives_nan_regions(gpu_ctx, speckle):
    """A flat block in the reference: the regions inside it hit the reference's catch
    (constant template -> NaN, Phi 0), their children stay NaN, gridfit drops them from
    the u / v fit but keeps their Phi (u, v and Phi fitted apart)."""
    import mg_oracle as mg
    _flush("search_mg: 200 x 200 pair with a flat 60 x 60 block, ws 16, step 8")
    p = speckle.small_pair(n=200, u0=1.6, v0=0.7, seed=8)
    gpu_ctx.set_images(p.f, p.g)
    f, g, _, _ = gpu_ctx.get_images()
    f = f.copy()
    f[20:80, 30:90] = f[20, 30]
    gpu_ctx.set_normalized(f, g)
    d = gpu_ctx.integer_search_mg((1, 200, 1, 200), 16, 8)
    r = mg.search_mg(f, g, [1, 200], [1, 200], 16, 8)
    assert d["n_nan"] > 0 and d["n_nan"] == int(np.isnan(r["tempu"]).sum())
    _compare(d, r)


def test_search_mg_rejects_constant_reference(gpu_ctx, pkg):
    _flush("search_mg: constant reference ROI")
    f = np.full((96, 96), 7.0)
    g = np.random.default_rng(0).normal(size=(96, 96))
    gpu_ctx.set_normalized(f, g)
    with pytest.raises(pkg.AldicError, match="constant"):
        gpu_ctx.integer_search_mg((1, 96, 1, 96), 16, 8)
