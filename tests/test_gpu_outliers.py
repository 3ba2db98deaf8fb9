"""The search's outlier test and NaN in-painting on the device against the oracle:
aldic_inpaint_nans (inpaint_nans.m methods 3 and 4), aldic_remove_outliers
(funRemoveOutliers' single-precision median test + springs), and the frame's
IntegerSearch/Init steps (mask holes, median test, in-painting) before LocalICGN."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _field(n, m, seed):
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:n, 0:m].astype(float)
    return np.sin(0.21 * r) + 0.05 * c + 0.002 * r * c + rng.normal(scale=0.01, size=(n, m))


PATTERNS = {
    "single_interior": lambda A: A.__setitem__((10, 11), np.nan),
    "cluster": lambda A: A.__setitem__((slice(12, 15), slice(7, 11)), np.nan),
    "first_row_and_corner": lambda A: (A.__setitem__((0, slice(3, 6)), np.nan),
                                       A.__setitem__((0, 0), np.nan)),
    "second_ring_and_last_column": lambda A: (A.__setitem__((1, slice(4, 9)), np.nan),
                                              A.__setitem__((slice(5, 9), -1), np.nan)),
    "large_hole": lambda A: A.__setitem__((slice(8, 20), slice(6, 18)), np.nan),
    "scattered": lambda A: A.__setitem__(np.unravel_index(
        np.random.default_rng(4).choice(A.size, A.size // 30, replace=False), A.shape), np.nan),
}


@pytest.mark.parametrize("pattern", list(PATTERNS))
@pytest.mark.parametrize("method", [3, 4])
def test_inpaint_matches_oracle(gpu_ctx, orc, pattern, method):
    _flush(f"inpaint {method} {pattern}")
    A = _field(29, 37, 1)
    PATTERNS[pattern](A)
    got = gpu_ctx.inpaint_nans(A, method)
    ref = orc.inpaint_plate(A) if method == 3 else orc.inpaint_springs(A)
    known = ~np.isnan(A)
    assert np.array_equal(got[known], A[known])
    # springs: a well-conditioned Laplace system; plate: normal equations of a del^4
    # least squares against the oracle's SVD solve
    tol = 1e-10 if method == 4 else (1e-6 if pattern == "large_hole" else 1e-8)
    assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref)))


def test_inpaint_errors(gpu_ctx):
    err = __import__("importlib").import_module("2d_aldic_amd").AldicError
    with pytest.raises(err):
        gpu_ctx.inpaint_nans(np.full((4, 5), np.nan), 4)
    with pytest.raises(err):
        gpu_ctx.inpaint_nans(np.zeros((4, 5)), 2)
    B = _field(6, 7, 2)
    assert np.array_equal(gpu_ctx.inpaint_nans(B, 3), B)   # nothing to fill


@pytest.mark.parametrize("thr", [100.0, 8.0, 2.0])
def test_remove_outliers_matches_oracle(gpu_ctx, orc, thr):
    n, m = 31, 40
    r, c = np.mgrid[0:n, 0:m].astype(float)
    rng = np.random.default_rng(int(thr))
    u = 2.0 + 0.01 * c - 0.02 * r + rng.normal(scale=0.05, size=(n, m))
    v = -1.0 + 0.015 * r + rng.normal(scale=0.05, size=(n, m))
    for (i, j, du) in ((4, 7, 40.0), (20, 33, -45.0), (30, 0, 30.0), (12, 39, 35.0)):
        u[i, j] += du
        v[i, j] -= 0.5 * du
    u[15:18, 20:23] = np.nan
    v[15:18, 20:23] = np.nan
    u[0, 10] = v[0, 10] = np.nan
    got_u, got_v, got_f = gpu_ctx.remove_outliers(u, v, thr)
    ref_u, ref_v, ref_f = orc.remove_outliers(u, v, thr)
    assert np.array_equal(got_f, ref_f)
    assert ref_f.sum() >= 3
    assert np.array_equal(np.isnan(got_u), np.isnan(u)) and np.array_equal(np.isnan(got_v), np.isnan(v))
    ok = ~np.isnan(u)
    assert np.max(np.abs(got_u[ok] - ref_u[ok])) < 1e-9
    assert np.max(np.abs(got_v[ok] - ref_v[ok])) < 1e-9


@pytest.mark.parametrize("thr", [100.0, 4.0])
def test_frame_outlier_stage_matches_oracle(gpu_ctx, speckle, orc, thr):
    """A decorrelated patch in the deformed image (wrong integer peaks), a flat patch in
    both images (constant templates: NaN) and a mask hole: the frame's search field goes
    through the mask, the median test and the in-painting before LocalICGN, as in the
    oracle's restatement of IntegerSearch.m and Init.m."""
    _flush(f"frame outlier stage thr {thr}")
    p = speckle.small_pair(n=256, u0=2.3, v0=-1.6, ux=0.002, vy=-0.001, bits=8, seed=12)
    f, g = p.f.copy(), p.g.copy()
    rng = np.random.default_rng(3)
    g[150:185, 60:95] = rng.integers(0, 255, size=(35, 35), dtype=np.uint8)
    f[40:75, 170:215] = 90
    g[40:75, 172:217] = 90
    mask = np.ones((256, 256), np.uint8)
    mask[100:122, 100:140] = 0
    grid = [1, 256, 1, 256]
    gpu_ctx.frame_stage(f, g, mask=mask)
    gpu_ctx.frame_setup(20, 10, 8, grid, tol=1e-2, mu=1e-3, beta=0.1, n_admm=3,
                        outlier_thr=thr)
    gpu_ctx.frame_run()
    out = gpu_ctx.frame_get()
    o = orc.frame(f, g, 20, 10, 8, grid, tol=1e-2, mu=1e-3, beta=0.1, n_admm=3, mask=mask,
                  outlier_thr=thr)
    # the initial guesses after the outlier stage
    assert not np.any(np.isnan(out["U0"]))
    assert np.max(np.abs(out["U0"] - o["U0"])) < 1e-9
    r = orc.search(*[orc.normalize(x.astype(float))[0] for x in (f, g)], grid, 20, 10, [8, 8])
    raw = np.stack([r["u"].ravel(), r["v"].ravel()], 1).ravel()
    changed = ~np.isclose(out["U0"], raw, equal_nan=False)
    assert changed.sum() > 0            # NaNs and (at 4) outliers were replaced
    assert np.array_equal(out["iters6"], o["iters6"])
    assert np.max(np.abs(out["U"] - o["U"])) < 1e-4
