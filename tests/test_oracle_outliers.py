"""The oracle's restatement of the in-painting methods (inpaint_nans.m 3 and 4) and of
funRemoveOutliers' median test, checked against what their definitions fix (the device
versions are compared with these in tests/test_gpu_outliers.py)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def al(orc):
    return orc


def test_inpaint_plate_reproduces_linear_and_quadratic_fields(al):
    # every stencil of method 3 (del^4 inside, the 5-point Laplacian on the second ring,
    # second differences on the border rows/columns) annihilates linear fields, and
    # del^4 annihilates quadratics, so holes are filled exactly
    n, m = 14, 17
    r, c = np.mgrid[0:n, 0:m].astype(float)
    lin = 0.3 + 0.7 * r - 1.1 * c
    A = lin.copy()
    A[5:8, 6:9] = np.nan     # interior hole
    A[0, 4] = np.nan         # on the first row
    A[9, m - 1] = np.nan     # on the last column
    A[1, 1] = np.nan         # second ring
    assert np.max(np.abs(al.inpaint_plate(A) - lin)) < 1e-9
    quad = 0.5 + 0.2 * r - 0.1 * c + 0.03 * r * r - 0.02 * r * c + 0.01 * c * c
    B = quad.copy()
    B[6:9, 7:10] = np.nan    # well inside: only del^4 rows touch it
    assert np.max(np.abs(al.inpaint_plate(B) - quad)) < 1e-9
    assert np.array_equal(al.inpaint_plate(lin), lin)


def test_inpaint_springs_isolated_nan_is_neighbour_mean(al):
    rng = np.random.default_rng(2)
    A = rng.normal(size=(9, 11))
    A[4, 5] = np.nan
    out = al.inpaint_springs(A)
    assert np.isclose(out[4, 5], np.mean([A[3, 5], A[5, 5], A[4, 4], A[4, 6]]), atol=1e-12)
    C = np.full((6, 7), 2.5)
    C[0, 0] = C[3, 3] = np.nan
    assert np.allclose(al.inpaint_springs(C), 2.5, atol=1e-12)


def test_median_filter_neighbourhood_orientation(al):
    # medFilt2 with a 2 x 2 window, padding after, and the 2nd member skipped: the median
    # of V(i, j), V(i+1, j), V(i+1, j+1).  A ramp along the columns is left unchanged, a
    # ramp along the rows moves up by one step except on the (mirrored) last row.
    n, m = 6, 8
    r, c = np.mgrid[0:n, 0:m].astype(float)
    assert np.array_equal(al.medfilt3(c), c)
    med = al.medfilt3(r)
    assert np.array_equal(med[:-1], r[:-1] + 1) and np.array_equal(med[-1], r[-1])
    # a single spike never reaches the median of any neighbourhood
    S = np.zeros((n, m))
    S[2, 3] = 100.0
    assert np.all(al.medfilt3(S) == 0.0)
    # values are taken in single precision
    x = np.full((3, 3), 1.0 + 1e-12)
    assert np.all(al.medfilt3(x) == np.float32(1.0 + 1e-12))


def test_remove_outliers_replaces_only_the_spike(al):
    n, m = 12, 15
    r, c = np.mgrid[0:n, 0:m].astype(float)
    u = 0.2 * r + 0.1 * c
    v = -0.05 * r + 0.3 * c
    u[5, 6] += 40.0          # an outlier in u
    u[9, 2] = np.nan         # a search failure stays NaN
    v[9, 2] = np.nan
    u2, v2, flagged = al.remove_outliers(u, v, 100)
    assert np.argwhere(flagged).tolist() == [[5, 6]]
    exp = np.mean([u[4, 6], u[6, 6], u[5, 5], u[5, 7]])
    assert np.isclose(u2[5, 6], exp, atol=1e-12)
    keep = np.ones((n, m), bool)
    keep[5, 6] = False
    keep[9, 2] = False
    assert np.array_equal(u2[keep], u[keep]) and np.array_equal(v2[keep], v[keep])
    assert np.isnan(u2[9, 2]) and np.isnan(v2[9, 2])


def test_median_test_is_single_precision(al):
    # fluct = single(u) - med: a step of 1e-9 on a level of 1 vanishes in single
    u = np.full((5, 6), 1.0)
    u[2, 3] += 1e-9
    _, nf = al.med_fluct(u)
    assert nf.dtype == np.float32 and np.all(nf == 0)


def test_remove_outliers_then_springs_equals_one_joint_solve(al):
    # funRemoveOutliers fills the flagged with the NaNs as unknowns, puts the NaNs back,
    # and Init.m:13 fills them with the flagged fixed: the same Laplace equations, so
    # one joint solve gives the same values
    rng = np.random.default_rng(9)
    n, m = 13, 16
    r, c = np.mgrid[0:n, 0:m].astype(float)
    u = np.sin(0.3 * r) + 0.1 * c
    v = np.cos(0.2 * c) - 0.05 * r
    u[4, 7] += 30.0
    v[8, 3] -= 25.0
    u[6:8, 10:12] = np.nan
    v[6:8, 10:12] = np.nan
    u2, v2, fl = al.remove_outliers(u, v, 100)
    assert fl[4, 7] and fl[8, 3]
    two = al.inpaint_springs(u2)
    joint = u.copy()
    joint[fl] = np.nan
    assert np.allclose(two, al.inpaint_springs(joint), atol=1e-12)
