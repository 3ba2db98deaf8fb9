"""The oracle's ComputeStrain restatement (ComputeStrain.m, funDerivativeOp.m,
PlaneFit2.m) pinned by independent formulations: numpy's first-order-edge gradient, exact
derivatives of linear and quadratic fields, numpy's least-squares solver, the matrix form
of the Green-Lagrangian strain, and the index pattern of MATLAB's column-major reshape."""
import numpy as np
import pytest


def _grid(U, M, N):
    return U[0::2].reshape(N, M).T, U[1::2].reshape(N, M).T   # [i, j], i along x


def test_derivative_op_is_central_inside_one_sided_on_the_border(orc):
    M, N, h = 7, 5, 3.0
    rng = np.random.default_rng(1)
    U = rng.normal(size=2 * M * N)
    F = orc.derivative_op(M, N, h) @ U
    u, v = _grid(U, M, N)
    # numpy: central differences inside, (f1 - f0) / h at the ends
    for comp, (field, axis) in enumerate([(u, 0), (v, 0), (u, 1), (v, 1)]):
        ref = np.gradient(field, h, axis=axis, edge_order=1)
        assert np.allclose(F[comp::4].reshape(N, M).T, ref, rtol=0, atol=1e-13)


def test_derivative_op_single_line_has_no_entries(orc):
    A = orc.derivative_op(1, 4, 2.0)
    U = np.full(8, np.nan)
    U[1::2] = 1.0
    F = A @ np.nan_to_num(U)
    assert np.all(F[0::4] == 0) and np.all(F[1::4] == 0)   # x rows: duplicates sum to 0
    assert A[0::4].nnz == 0 and A[1::4].nnz == 0


def test_linear_field_exact_everywhere(orc):
    M, N, h = 9, 6, 4.0
    i, j = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    x, y = h * i, h * j
    u = 0.3 + 0.01 * x - 0.02 * y
    v = -1.0 + 0.005 * x + 0.03 * y
    U = np.empty(2 * M * N)
    U[0::2] = u.flatten(order="F")
    U[1::2] = v.flatten(order="F")
    for method, rad in ((1, 1), (2, 1), (2, 2)):
        out = orc.compute_strain(M, N, h, U=U, method=method, rad=rad)
        F = out["FStrain"]
        assert np.allclose(F[0::4], 0.01, atol=1e-14) and np.allclose(F[1::4], 0.005, atol=1e-14)
        assert np.allclose(F[2::4], -0.02, atol=1e-14) and np.allclose(F[3::4], 0.03, atol=1e-14)


def test_plane_fit_matches_lstsq_and_quadratic_slopes(orc):
    rng = np.random.default_rng(2)
    M, N, R, h = 8, 7, 2, 1.5
    U = rng.normal(size=(M, N))
    UY, UX, UNew = orc.plane_fit2(U, h, h, R)
    for ii in range(R, M - R):
        for jj in range(R, N - R):
            A, b = [], []
            for tj in range(-R, R + 1):
                for ti in range(-R, R + 1):
                    A.append([1, ti * h, tj * h])
                    b.append(U[ii + ti, jj + tj])
            c = np.linalg.lstsq(np.array(A), np.array(b), rcond=None)[0]
            assert np.allclose([UNew[ii, jj], UX[ii, jj], UY[ii, jj]], c, atol=1e-13)
    # a quadratic on the symmetric stencil: the fitted slope is the derivative at the node
    i, j = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    Q = 0.2 * (h * i) ** 2 - 0.1 * (h * i) * (h * j) + 0.05 * (h * j) ** 2
    UY, UX, _ = orc.plane_fit2(Q, h, h, 1)
    dx = 0.4 * h * i - 0.1 * h * j
    dy = -0.1 * h * i + 0.1 * h * j
    assert np.allclose(UX[1:-1, 1:-1], dx[1:-1, 1:-1], atol=1e-12)
    assert np.allclose(UY[1:-1, 1:-1], dy[1:-1, 1:-1], atol=1e-12)
    assert np.array_equal(UX[0], Q[0])        # the border keeps the input


def test_strain_types(orc):
    rng = np.random.default_rng(3)
    Fs = rng.normal(scale=0.05, size=40)
    E0 = orc.strain_type(Fs, 0)
    assert np.array_equal(E0, Fs)
    E2 = orc.strain_type(Fs, 2)
    E1 = orc.strain_type(Fs, 1)
    for n in range(10):
        ux, vx, uy, vy = Fs[4 * n:4 * n + 4]
        G = np.array([[ux, uy], [vx, vy]])
        A = np.eye(2) - G
        E = 0.5 * (np.eye(2) - A.T @ A)                # E = (C - I) / 2 with F = I - grad u
        assert np.allclose(E2[4 * n:4 * n + 4], [E[0, 0], E[1, 0], E[0, 1], E[1, 1]], atol=1e-15)
        assert np.isclose(E1[4 * n], ux / (1 - ux)) and np.isclose(E1[4 * n + 3], vy / (1 - vy))
        assert np.isclose(E1[4 * n + 2], uy / (1 - vy)) and np.isclose(E1[4 * n + 1], vx / (1 - ux))
    with pytest.raises(ValueError):
        orc.strain_type(Fs, 3)


def test_trim_index_and_world_flip(orc):
    M, N, r = 6, 5, 1
    t3 = orc.trim_index(M, N, r)
    want = []
    for j in range(r, N - r):
        for i in range(r, M - r):
            n = i + M * j
            want += [4 * n, 4 * n + 1, 4 * n + 2, 4 * n + 3]
    assert np.array_equal(t3, want)
    rng = np.random.default_rng(4)
    F = rng.normal(size=4 * M * N)
    out = orc.compute_strain(M, N, 2.0, F=F, method=0)
    assert out["Rad"] == 0 and np.array_equal(out["FStraintemp"], F)
    W = out["FStrainWorld"]
    assert np.array_equal(W[0::4], F[0::4]) and np.array_equal(W[1::4], -F[1::4])
    assert np.array_equal(W[2::4], -F[2::4]) and np.array_equal(W[3::4], F[3::4])


def test_plane_fit_method_keeps_differences_on_the_ring(orc):
    rng = np.random.default_rng(5)
    M, N, h, R = 9, 8, 2.0, 2
    U = rng.normal(size=2 * M * N)
    fd = orc.compute_strain(M, N, h, U=U, method=1)["FStrain"]
    pf = orc.compute_strain(M, N, h, U=U, method=2, rad=R)
    inner = np.zeros(4 * M * N, bool)
    inner[orc.trim_index(M, N, R)] = True
    assert np.array_equal(pf["FStrain"][~inner], fd[~inner])
    assert not np.allclose(pf["FStrain"][inner], fd[inner])
    assert pf["Rad"] == R and pf["FStrainWorld"].size == 4 * (M - 2 * R) * (N - 2 * R)
