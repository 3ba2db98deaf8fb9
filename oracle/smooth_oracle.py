"""CPU restatement of the quadtree strain smoothing, for the tests only:
funSmoothStrainQuadtree (func_quadtree/funSmoothStrainQuadtree.m:27-104) and the
regularizeNd it calls (func/regularizeNd.m:1-937, 2-D, 'linear' interpolation, the
'normal' solver: its defaults when called with four arguments).

regularizeNd fits a grid to scattered data by least squares of the stacked system
[A_fidelity; L_1; L_2] z = [y; 0] (regularizeNd.m:679-720, solved through the normal
equations), where A_fidelity interpolates the grid bilinearly at the data points
(:384-587) and L_d takes second derivatives along dimension d at every interior grid
point (:588-677), scaled by smoothness * sqrt(n_points / n_equations_d) *
(grid range_d)^2 with the three-point weights 2 / ((x1 - x3)(x1 - x2)), ... of
secondDerivativeWeights (:766-800).
"""
import numpy as np

import oracle


def _cells(grid, v):
    """cell index (0-based) and local coordinate of each value, the last grid point in the
    last cell (regularizeNd.m:436-460: histc with the top edge folded in)"""
    grid = np.asarray(grid, np.float64)
    k = np.searchsorted(grid, v, side="right") - 1
    k = np.clip(k, 0, grid.size - 2)
    t = (v - grid[k]) / (grid[k + 1] - grid[k])
    return k, t


def regularize_nd_system(x, y, z, xg, yg, smoothness):
    """(A, rhs) of the stacked least-squares system; grid index ix + nx * iy (dimension 1
    fastest, as MATLAB's column-major (nx, ny) yGrid)."""
    import scipy.sparse as sp
    x = np.asarray(x, np.float64).ravel()
    y = np.asarray(y, np.float64).ravel()
    z = np.asarray(z, np.float64).ravel()
    xg = np.asarray(xg, np.float64).ravel()
    yg = np.asarray(yg, np.float64).ravel()
    nx, ny, n = xg.size, yg.size, x.size
    kx, tx = _cells(xg, x)
    ky, ty = _cells(yg, y)
    rows = np.repeat(np.arange(n), 4)
    cols = np.stack([kx + nx * ky, kx + 1 + nx * ky, kx + nx * (ky + 1), kx + 1 + nx * (ky + 1)], 1).ravel()
    vals = np.stack([(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty], 1).ravel()
    blocks = [sp.csr_matrix((vals, (rows, cols)), shape=(n, nx * ny))]
    s = np.broadcast_to(np.asarray(smoothness, np.float64), (2,))
    for d, (g, m, other) in enumerate(((xg, nx, ny), (yg, ny, nx))):
        if s[d] == 0:
            continue
        neq = (m - 2) * other
        x1, x2, x3 = g[:-2], g[1:-1], g[2:]
        w = np.stack([2 / ((x1 - x3) * (x1 - x2)), 2 / ((x2 - x1) * (x2 - x3)), 2 / ((x3 - x1) * (x3 - x2))], 1)
        scale = s[d] * np.sqrt(n / neq) * (g[-1] - g[0]) ** 2
        i = np.arange(1, m - 1)
        r, c, v = [], [], []
        for o in range(other):
            for a in range(3):
                idx = i - 1 + a
                gi = idx + nx * o if d == 0 else o + nx * idx
                r.append(np.arange(m - 2) + (m - 2) * o)
                c.append(gi)
                v.append(scale * w[:, a])
        blocks.append(sp.csr_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))),
                                    shape=(neq, nx * ny)))
    A = sp.vstack(blocks).tocsr()
    rhs = np.concatenate([z, np.zeros(A.shape[0] - n)])
    return A, rhs


def regularize_nd(x, y, z, xg, yg, smoothness):
    """regularizeNd([x y], z, {xg, yg}, smoothness): the (nx, ny) grid ([ix, iy])."""
    import scipy.sparse.linalg as spl
    A, rhs = regularize_nd_system(x, y, z, xg, yg, smoothness)
    N = (A.T @ A).tocsc()
    zz = spl.spsolve(N, A.T @ rhs)
    return zz.reshape(len(yg), len(xg)).T


def smooth_strain_quadtree(F, coords, h, smoothness=1e-4, sigma=0.5, size=3):
    """funSmoothStrainQuadtree(F, DICmesh, DICpara) with elementMinSize h,
    StrainSmoothness smoothness, StrainFilterStd sigma and StrainFilterSize size: two
    passes (SmoothTimes 1, 2) of: every component regularizeNd'ed over the grid
    min:h:max in x and y, blurred by nanconv(., fspecial('gaussian', size, sigma), 'edge',
    'nanout') and read back at the nodes (which lie on the grid)."""
    coords = np.asarray(coords, np.float64).reshape(-1, 2)
    F = np.asarray(F, np.float64).reshape(-1).copy()
    xg = np.arange(coords[:, 0].min(), coords[:, 0].max() + 1e-9, h)
    yg = np.arange(coords[:, 1].min(), coords[:, 1].max() + 1e-9, h)
    ix = np.searchsorted(xg, coords[:, 0])
    iy = np.searchsorted(yg, coords[:, 1])
    if not (np.all(ix < xg.size) and np.all(xg[np.minimum(ix, xg.size - 1)] == coords[:, 0]) and
            np.all(iy < yg.size) and np.all(yg[np.minimum(iy, yg.size - 1)] == coords[:, 1])):
        raise ValueError("a node is off the smoothing grid (the reference's find fails)")
    k = oracle.gauss_kernel(size, sigma)
    for _ in range(2):
        out = F.copy()
        for c in range(4):
            Z = regularize_nd(coords[:, 0], coords[:, 1], F[c::4], xg, yg, smoothness)
            B = oracle.nanconv_edge(Z, k)
            out[c::4] = B[ix, iy]
        F = out
    return F
