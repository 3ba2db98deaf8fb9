"""The CPU restatement of funIntegerSearchMg (oracle/mg_oracle.py), pinned on its own:
its rectangular normxcorr2 against the oracle's square one and a brute-force loop, the
cropped lags against the full map, the recursion's level and region counts against a
hand count, and a pure whole-pixel translation recovered at every region and grid node."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mg_oracle as mg  # noqa: E402
import oracle  # noqa: E402


def _brute(T, A):
    m, n = T.shape
    M, N = A.shape
    Tp = T - T.mean()
    st = np.sum(Tp * Tp)
    out = np.zeros((m + M - 1, n + N - 1))
    Ap = np.zeros((M + 2 * m, N + 2 * n))
    Ap[m:m + M, n:n + N] = A
    for r in range(m + M - 1):
        for c in range(n + N - 1):
            W = Ap[r + 1:r + 1 + m, c + 1:c + 1 + n]
            va = max(np.sum(W * W) - W.sum() ** 2 / (m * n), 0.0)
            den = np.sqrt(st * va)
            out[r, c] = np.sum(Tp * W) / den if den > 0 else 0.0
    return out


def test_ncc_full_matches_square_oracle_and_brute_force():
    rng = np.random.default_rng(0)
    T = rng.normal(size=(7, 7))
    A = rng.normal(size=(12, 12))
    full = mg.ncc_full(T, A)
    # the oracle's square-template map, stored y-major: transpose to [x, y]
    sq, ok = oracle.ncc_full(T.T, A.T)
    assert ok == 0
    assert np.max(np.abs(full - sq.T)) < 1e-12
    T2 = rng.normal(size=(5, 8))
    A2 = rng.normal(size=(9, 11))
    A2[0:3, :] = 0.0                       # zero-variance windows at the edge lags
    assert np.max(np.abs(mg.ncc_full(T2, A2) - _brute(T2, A2))) < 1e-12


def test_ncc_crop_equals_full_map_entries():
    rng = np.random.default_rng(1)
    T = rng.normal(size=(9, 6))
    A = rng.normal(size=(19, 16))
    full = mg.ncc_full(T, A)
    ox, oy = np.arange(-1, 12), np.arange(-1, 13)
    crop = mg.ncc_crop(T, A, ox, oy)
    m, n = T.shape
    assert np.max(np.abs(crop - full[np.ix_(ox + m - 1, oy + n - 1)])) < 1e-12


def _speckle(n, seed):
    rng = np.random.default_rng(seed)
    img = np.zeros((n, n))
    y, x = np.mgrid[0:n, 0:n]
    for cx, cy, a in zip(rng.uniform(0, n, 900), rng.uniform(0, n, 900), rng.uniform(0.4, 1, 900)):
        img += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * 2.0 ** 2))
    return img


def test_whole_pixel_translation_and_level_counts():
    n, du, dv = 200, 3, -2
    big = _speckle(n + 20, 4)
    f = big[10:10 + n, 10:10 + n]
    g = big[10 - dv:10 - dv + n, 10 - du:10 - du + n]     # g(x + u) = f(x)
    r = mg.search_mg(f, g, [1, n], [1, n], 32, 16)
    # ROI 25..175 (gap 1 + round(24) = 25): width 150 -> 75 -> 37.5 < 32? no, 37.5 -> 18.75:
    # levels 1 (whole ROI), 2 (4 regions), 3 (16), 4 (64 regions of 18-19 px)
    assert r["levels"] == 4 and r["tempx"].size == 64
    # every region peaks at the shift (findpeak's paraboloid moves it by its usual bias,
    # up to ~0.25 px on the ~19 px regions of the last level)
    assert np.all(np.rint(r["tempu"]) == du) and np.all(np.rint(r["tempv"]) == dv)
    assert np.all(np.abs(r["tempu"] - du) < 0.3) and np.all(np.abs(r["tempv"] - dv) < 0.3)
    # the springs fit extrapolates at the grid corners, beyond the region centres
    assert np.all(np.abs(r["u"] - du) < 0.3) and np.all(np.abs(r["v"] - dv) < 0.3)
    assert np.mean(np.abs(r["u"] - du)) < 0.05 and np.mean(np.abs(r["v"] - dv)) < 0.05
    assert r["u"].shape == (r["ylist"].size, r["xlist"].size)
    assert np.all(r["ccmax"] > 0.9)
