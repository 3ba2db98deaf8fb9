"""ctypes front end of the FP64 CPU oracle (liborc.so, built from aldic_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
the CPU-baseline leg of ``bench.py``, where it is the reference the HIP library is
compared against and the CPU baseline.
The product (the HIP library and its Python mirror) never imports this module.

Array conventions follow aldic_oracle.c: images (ny, nx) C-ordered (x fastest),
1-based pixel coordinates, interleaved nodal vectors.

Pinning of the functions restated here (tests/test_oracle_pinning.py):
  gridfit_system/gridfit  springs energy against an explicit edge loop, |Areg|_1 closed
                          forms, dense QR least squares, constants, coordinate scaling
  subpb2_quadtree         equals the Q4 step without hanging nodes; patch test with
                          hanging and Dirichlet nodes
  gauss_strain_quadtree   bilinear fields give the exact gradient at the Gauss points,
                          whose positions and order are checked in closed form
  gather_weights          weight rows that select single points and linear combinations
  badpoint                hand-counted thresholds of rules 0, 1 and 2
  sibson, natural_fill    1/4 weights of an isolated lattice hole, weights >= 0 summing to
                          1 with linear precision, brute-force area sampling, the hull
                          rule's closed forms (tests/test_oracle_natural.py)
  gauss_kernel, nanconv_edge, smooth_nodal, beta_choice
                          fspecial / scipy.ndimage equivalents, brute force, the
                          parabola vertex
  frame                   composed of the above
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build(force=False):
    so = os.path.join(_HERE, "liborc.so")
    src = os.path.join(_HERE, "aldic_oracle.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-B" if force else "liborc.so"], cwd=_HERE)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        I, D, V = C.c_int, C.c_double, None
        L.orc_normalize.argtypes = [f64p, I, I, i32p, f64p, f64p]
        L.orc_gradient.argtypes = [f64p, I, I, C.c_void_p, f64p, f64p]
        L.orc_bicubic.argtypes = [f64p, I, I, f64p, f64p, C.c_long, f64p]
        L.orc_ic_compose.argtypes = [f64p, f64p, C.POINTER(I)]
        L.orc_icgn6.argtypes = [f64p, f64p, f64p, f64p, I, I, I, f64p, f64p, i32p, D,
                                f64p, f64p, f64p, i32p, f64p]
        L.orc_icgn2.argtypes = [f64p, f64p, f64p, f64p, I, I, I, f64p, i32p, f64p, D,
                                f64p, f64p, f64p, f64p, D, f64p, i32p]
        L.orc_search_grid.argtypes = [I, I, i32p, I, I, i32p, i32p]
        L.orc_ncc_full.argtypes = [f64p, I, f64p, I, I, f64p, C.POINTER(I)]
        L.orc_findpeak.argtypes = [f64p, I, I, f64p]
        L.orc_search.argtypes = [f64p, f64p, I, I, I, I, I, I, I, I, i32p, f64p, f64p, f64p,
                                 f64p, f64p, f64p, f64p, i32p]
        L.orc_search_sample.argtypes = [f64p, f64p, I, I, I, I, I, I, I, i32p, I, i32p, f64p,
                                        f64p, f64p, f64p, f64p, i32p]
        L.orc_subpb2_assemble.argtypes = [f64p, I, i32p, I, D, D, D, f64p, f64p, f64p, f64p,
                                          i32p, i32p, f64p, f64p]
        L.orc_pcg.argtypes = [I, i32p, i32p, f64p, f64p, f64p, D, I]
        L.orc_pcg.restype = I
        L.orc_nodal_strain.argtypes = [f64p, I, i32p, I, f64p, I, f64p]
        L.orc_sibson.argtypes = [I, f64p, D, D, i32p, f64p, I]
        L.orc_sibson.restype = I
        L.orc_natural_fill.argtypes = [I, f64p, i32p, I, f64p, i32p]
        L.orc_natural_fill.restype = I
        for fn in ("orc_normalize", "orc_gradient", "orc_bicubic", "orc_ic_compose", "orc_icgn6",
                   "orc_icgn2", "orc_search_grid", "orc_ncc_full", "orc_findpeak", "orc_search",
                   "orc_search_sample",
                   "orc_subpb2_assemble", "orc_nodal_strain"):
            getattr(L, fn).restype = V
        _LIB = L
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def normalize(img, roi=None):
    img = _f64(img)
    ny, nx = img.shape
    roi = _i32(roi if roi is not None else [1, nx, 1, ny])
    out = np.empty_like(img)
    ms = np.zeros(2)
    lib().orc_normalize(img, nx, ny, roi, out, ms)
    return out, ms


def gradient(f, mask=None):
    f = _f64(f)
    ny, nx = f.shape
    gx = np.empty_like(f)
    gy = np.empty_like(f)
    m = None
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_gradient(f, nx, ny, m.ctypes.data if m is not None else None, gx, gy)
    return gx, gy


def bicubic(img, X, Y):
    img = _f64(img)
    ny, nx = img.shape
    X = _f64(X).ravel()
    Y = _f64(Y).ravel()
    Z = np.empty_like(X)
    lib().orc_bicubic(img, nx, ny, X, Y, X.size, Z)
    return Z


def ic_compose(P, D):
    P = _f64(P).copy()
    ok = C.c_int(0)
    lib().orc_ic_compose(P, _f64(D), C.byref(ok))
    return P, ok.value


def _ws(ws, S):
    if np.isscalar(ws):
        return np.full(S, int(ws), np.int32)
    return _i32(ws)


def icgn6(f, g, gx, gy, coords, U0, ws, tol=1e-2):
    f, g, gx, gy = map(_f64, (f, g, gx, gy))
    ny, nx = f.shape
    coords = _f64(coords).reshape(-1, 2)
    S = coords.shape[0]
    U = np.empty(2 * S)
    F = np.empty(4 * S)
    H = np.empty((S, 21))
    it = np.empty(S, np.int32)
    z = np.empty(S)
    lib().orc_icgn6(f, g, gx, gy, nx, ny, S, coords, _f64(U0), _ws(ws, S), tol, U, F, H, it, z)
    return U, F, H, it, z


def icgn2(f, g, gx, gy, coords, ws, H21, mu, udual, vdual, Uold, Fold, tol=1e-2):
    f, g, gx, gy = map(_f64, (f, g, gx, gy))
    ny, nx = f.shape
    coords = _f64(coords).reshape(-1, 2)
    S = coords.shape[0]
    U = np.empty(2 * S)
    it = np.empty(S, np.int32)
    lib().orc_icgn2(f, g, gx, gy, nx, ny, S, coords, _ws(ws, S), _f64(H21), mu, _f64(udual),
                    _f64(vdual), _f64(Uold), _f64(Fold), tol, U, it)
    return U, it


def search_grid(nx, ny, grid, w, step, s):
    out = np.zeros(4, np.int32)
    lib().orc_search_grid(nx, ny, _i32(grid), w, step, _i32(s), out)
    return [int(v) for v in out]


def ncc_full(T, D):
    T = _f64(T)
    D = _f64(D)
    tw = T.shape[0]
    Dy, Dx = D.shape
    Cm = np.empty((tw - 1 + Dy, tw - 1 + Dx))
    ok = C.c_int(0)
    lib().orc_ncc_full(T, tw, D, Dx, Dy, Cm, C.byref(ok))
    return Cm, ok.value


def findpeak(fmat):
    """findpeak on a MATLAB-shaped (nr, nc) matrix; returns (xpeak=col, ypeak=row, max)."""
    fmat = np.asarray(fmat, dtype=np.float64)
    nr, nc = fmat.shape
    col_major = np.ascontiguousarray(fmat.T)  # dim 1 (rows) fastest
    out = np.zeros(3)
    lib().orc_findpeak(col_major, nr, nc, out)
    return out[0], out[1], out[2]


def search(f, g, grid, w, step, s):
    f = _f64(f)
    g = _f64(g)
    ny, nx = f.shape
    s = _i32(s if np.ndim(s) else [s, s])
    x1, y1, Nx, Ny = search_grid(nx, ny, grid, w, step, s)
    n = Nx * Ny
    outs = [np.empty(n) for _ in range(7)]
    st = np.empty(n, np.int32)
    lib().orc_search(f, g, nx, ny, x1, y1, Nx, Ny, w, step, s, *outs, st)
    x, y, u, v, cm, pce, ppe = [o.reshape(Ny, Nx) for o in outs]
    return dict(x=x, y=y, u=u, v=v, ccmax=cm, pce=pce, ppe=ppe, status=st.reshape(Ny, Nx),
                x1=x1, y1=y1, Nx=Nx, Ny=Ny)


def search_sample(f, g, grid, w, step, s, qlist):
    """search() on the listed subsets only (entries l*Nx + k of the whole-field grid),
    cropped lags only, no q-factors -> per-entry x, y, u, v, ccmax, status."""
    f = _f64(f)
    g = _f64(g)
    ny, nx = f.shape
    s = _i32(s if np.ndim(s) else [s, s])
    x1, y1, Nx, Ny = search_grid(nx, ny, grid, w, step, s)
    q = _i32(qlist)
    assert q.size == 0 or (q.min() >= 0 and q.max() < Nx * Ny)
    n = q.size
    outs = [np.empty(n) for _ in range(5)]
    st = np.empty(n, np.int32)
    lib().orc_search_sample(f, g, nx, ny, x1, y1, Nx, w, step, s, n, q, *outs, st)
    x, y, u, v, cm = outs
    return dict(x=x, y=y, u=u, v=v, ccmax=cm, status=st)


def subpb2_assemble(coords, elems, beta, mu, U, F, udual, vdual, alpha=0.0):
    coords = _f64(coords).reshape(-1, 2)
    elems = _i32(elems).reshape(-1, 4)
    nn = coords.shape[0]
    E = elems.shape[0]
    ri = np.empty(E * 64, np.int32)
    ci = np.empty(E * 64, np.int32)
    val = np.empty(E * 64)
    b = np.empty(2 * nn)
    lib().orc_subpb2_assemble(coords, nn, elems, E, beta, mu, alpha, _f64(U), _f64(F),
                              _f64(udual), _f64(vdual), ri, ci, val, b)
    return ri, ci, val, b


def subpb2(coords, elems, beta, mu, U, F, udual, vdual, alpha=0.0, rtol=1e-14, maxit=20000):
    """Subpb2.m: assemble (sparse() sums the triplets) and solve in FP64."""
    import scipy.sparse as sp
    ri, ci, val, b = subpb2_assemble(coords, elems, beta, mu, U, F, udual, vdual, alpha)
    n = b.size
    A = sp.csr_matrix((val, (ri, ci)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    x = np.zeros(n)
    it = lib().orc_pcg(n, _i32(A.indptr), _i32(A.indices), _f64(A.data), b, x, rtol, maxit)
    return x, it


def nodal_strain(coords, elems, U, swap_fix=False):
    coords = _f64(coords).reshape(-1, 2)
    elems = _i32(elems).reshape(-1, 4)
    nn = coords.shape[0]
    F = np.empty(4 * nn)
    lib().orc_nodal_strain(coords, nn, elems, elems.shape[0], _f64(U), int(swap_fix), F)
    return F


# ---- gridfit (func/gridfit.m), the options Subpb1 uses ------------------------------

def gridfit_system(x, y, z, xnodes, ynodes, smoothness=1.0):
    """Least-squares system of gridfit.m with interp 'triangle', regularizer 'springs',
    autoscale on, no mask, no tiling (gridfit.m:295-306 defaults; Subpb1.m:188-189 sets
    'regularizer','springs').  Returns (A, Areg, rhs, lam) with the fit minimising
    |A z - rhs|^2 + lam^2 |Areg z|^2, lam = smoothness * |A|_1 / |Areg|_1 (:651-654).
    Grid index of node (ix, iy) is iy + ny * ix (MATLAB column-major (ny, nx))."""
    import scipy.sparse as sp
    x = np.asarray(x, np.float64).ravel()
    y = np.asarray(y, np.float64).ravel()
    z = np.asarray(z, np.float64).ravel()
    xnodes = np.asarray(xnodes, np.float64).ravel()
    ynodes = np.asarray(ynodes, np.float64).ravel()
    nx, ny = xnodes.size, ynodes.size
    dx, dy = np.diff(xnodes), np.diff(ynodes)
    xscale, yscale = dx.mean(), dy.mean()                       # :361-365
    n = x.size
    # :460-469 cell of each point (histc), the last node counted in the last cell
    indx = np.searchsorted(xnodes, x, side="right")             # 1-based histc bin
    indy = np.searchsorted(ynodes, y, side="right")
    indx[indx == nx] = nx - 1
    indy[indy == ny] = ny - 1
    ind = (indy - 1) + ny * (indx - 1)                          # 0-based grid index
    # :496-511 triangle interpolation inside the cell
    tx = np.clip((x - xnodes[indx - 1]) / dx[indx - 1], 0.0, 1.0)
    ty = np.clip((y - ynodes[indy - 1]) / dy[indy - 1], 0.0, 1.0)
    L = np.where(tx > ty, ny, 1)
    t1, t2 = np.minimum(tx, ty), np.maximum(tx, ty)
    rows = np.repeat(np.arange(n), 3)
    cols = np.stack([ind, ind + ny + 1, ind + L], 1).ravel()
    vals = np.stack([1 - t2, t1, t2 - t1], 1).ravel()
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, nx * ny))
    # :540-569 zero-rest-length springs: y neighbours, x neighbours, both diagonals
    blocks = []
    i, j = np.meshgrid(np.arange(nx), np.arange(ny - 1))        # ind = j + ny i
    ii, jj = i.ravel(), j.ravel()
    k = jj + ny * ii
    st = 1.0 / (dy[jj] / yscale)
    blocks.append((k, k + 1, st))
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny))
    ii, jj = i.ravel(), j.ravel()
    k = jj + ny * ii
    st = 1.0 / (dx[ii] / xscale)
    blocks.append((k, k + ny, st))
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1))
    ii, jj = i.ravel(), j.ravel()
    k = jj + ny * ii
    st = 1.0 / np.sqrt((dx[ii] / xscale) ** 2 + (dy[jj] / yscale) ** 2)
    blocks.append((k, k + ny + 1, st))
    blocks.append((k + 1, k + ny, st))
    r0, rr, cc, vv = 0, [], [], []
    for a, b, st in blocks:
        m = a.size
        rr += [np.arange(r0, r0 + m)] * 2
        cc += [a, b]
        vv += [-st, st]
        r0 += m
    Areg = sp.csr_matrix((np.concatenate(vv), (np.concatenate(rr), np.concatenate(cc))),
                         shape=(r0, nx * ny))
    NA = abs(A).sum(axis=0).max()                               # norm(A,1)
    NR = abs(Areg).sum(axis=0).max()
    return A, Areg, z, smoothness * NA / NR


def gridfit(x, y, z, xnodes, ynodes, smoothness=1.0):
    """gridfit.m (Subpb1's options): zgrid (ny x nx).  MATLAB's backslash gives the
    least-squares solution of [A; lam Areg] z = [rhs; 0]; here through the normal
    equations (the regularised system is well conditioned), FP64 sparse direct."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    A, Areg, rhs, lam = gridfit_system(x, y, z, xnodes, ynodes, smoothness)
    N = (A.T @ A + lam * lam * (Areg.T @ Areg)).tocsc()
    zz = spl.spsolve(N, A.T @ rhs)
    return zz.reshape(np.asarray(xnodes).size, np.asarray(ynodes).size).T


def subpb1_smooth(U, coords):
    """Subpb1.m:185-199: gridfit ('springs') of u and v over the nodes whose u is not
    NaN, evaluated back at every node."""
    coords = np.asarray(coords, np.float64).reshape(-1, 2)
    U = np.asarray(U, np.float64).copy()
    xn, yn = np.unique(coords[:, 0]), np.unique(coords[:, 1])
    ok = ~np.isnan(U[0::2])
    ix = np.searchsorted(xn, coords[:, 0])
    iy = np.searchsorted(yn, coords[:, 1])
    zu = gridfit(coords[ok, 0], coords[ok, 1], U[0::2][ok], xn, yn)
    zv = gridfit(coords[ok, 0], coords[ok, 1], U[1::2][ok], xn, yn)
    U[0::2] = zu[iy, ix]
    U[1::2] = zv[iy, ix]
    return U


def _q8_shape(ksi, eta, d):
    """Shape functions and their (ksi, eta) derivatives of the quadtree element with
    optional midside nodes 5..8 (Subpb2Quadtree.m:126-133 and its funDN* helpers):
    N5..N8 = d_k/2 (1 +- ksi or eta)(1 - |other|) on the right, top, left, bottom edges,
    corners bilinear minus half of their two neighbouring midside functions."""
    d5, d6, d7, d8 = d
    N5 = d5 * 0.5 * (1 + ksi) * (1 - abs(eta))
    N6 = d6 * 0.5 * (1 + eta) * (1 - abs(ksi))
    N7 = d7 * 0.5 * (1 - ksi) * (1 - abs(eta))
    N8 = d8 * 0.5 * (1 - eta) * (1 - abs(ksi))
    N = np.array([(1 - ksi) * (1 - eta) / 4 - 0.5 * (N7 + N8), (1 + ksi) * (1 - eta) / 4 - 0.5 * (N8 + N5),
                  (1 + ksi) * (1 + eta) / 4 - 0.5 * (N5 + N6), (1 - ksi) * (1 + eta) / 4 - 0.5 * (N6 + N7),
                  N5, N6, N7, N8])
    se, sk = np.sign(-eta), np.sign(-ksi)
    k5, e5 = d5 * 0.5 * (1 - abs(eta)), d5 * 0.5 * (1 + ksi) * se
    k6, e6 = d6 * 0.5 * (1 + eta) * sk, d6 * 0.5 * (1 - abs(ksi))
    k7, e7 = -d7 * 0.5 * (1 - abs(eta)), d7 * 0.5 * (1 - ksi) * se
    k8, e8 = d8 * 0.5 * (1 - eta) * sk, -d8 * 0.5 * (1 - abs(ksi))
    dk = np.array([-0.25 * (1 - eta) - 0.5 * (k7 + k8), 0.25 * (1 - eta) - 0.5 * (k8 + k5),
                   0.25 * (1 + eta) - 0.5 * (k5 + k6), -0.25 * (1 + eta) - 0.5 * (k6 + k7),
                   k5, k6, k7, k8])
    de = np.array([-0.25 * (1 - ksi) - 0.5 * (e7 + e8), -0.25 * (1 + ksi) - 0.5 * (e8 + e5),
                   0.25 * (1 + ksi) - 0.5 * (e5 + e6), 0.25 * (1 - ksi) - 0.5 * (e6 + e7),
                   e5, e6, e7, e8])
    return N, dk, de


def subpb2_quadtree(coords, elems8, dirichlet, beta, mu, U, F):
    """Subpb2Quadtree.m:1-346 with alpha = 0 and 2x2 Gauss points (as
    main_ALDIC_Quadtree.m:181-212 calls it): elements (E x 8, 0-based, -1 = no midside
    node); the duals are multiplied by 0 (:42-43); A = sum wJ (beta DN'DN + mu N'N),
    b = sum wJ (beta diag(DN' F_node') + mu N'N U) with each DOF's own nodal F; U fixed
    at the Dirichlet nodes; the free DOFs are the nodes of the elements minus the first
    entry of their sorted list (:165-168, which assumes a 0 there, i.e. that some element
    lacks a midside node) minus the Dirichlet nodes; every other node ends at 0."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    coords = np.asarray(coords, np.float64).reshape(-1, 2)
    el = np.asarray(elems8, np.int64).reshape(-1, 8)
    U = np.asarray(U, np.float64)
    F = np.asarray(F, np.float64)
    nn = coords.shape[0]
    g = 1.0 / np.sqrt(3.0)
    rows, cols, vals = [], [], []
    b = np.zeros(2 * nn)
    for e in range(el.shape[0]):
        nd = el[e]
        d = [1.0 if nd[k] >= 0 else 0.0 for k in range(4, 8)]
        px = np.array([coords[n, 0] if n >= 0 else 0.0 for n in nd])
        py = np.array([coords[n, 1] if n >= 0 else 0.0 for n in nd])
        Ae = np.zeros((16, 16))
        be = np.zeros(16)
        idx = np.array([2 * n + c if n >= 0 else -1 for n in nd for c in (0, 1)])
        for ksi in (-g, g):
            for eta in (-g, g):
                N, dk, de = _q8_shape(ksi, eta, d)
                J = np.array([[dk @ px, dk @ py], [de @ px, de @ py]])
                det = np.linalg.det(J)
                Ji = np.linalg.inv(J)
                dx = Ji[0, 0] * dk + Ji[0, 1] * de
                dy = Ji[1, 0] * dk + Ji[1, 1] * de
                Nm = np.zeros((2, 16)); Nm[0, 0::2] = N; Nm[1, 1::2] = N
                DN = np.zeros((4, 16)); DN[0, 0::2] = dx; DN[1, 0::2] = dy; DN[2, 1::2] = dx; DN[3, 1::2] = dy
                Ae += det * (beta * DN.T @ DN + mu * Nm.T @ Nm)
                Ue = np.array([U[i] if i >= 0 else 0.0 for i in idx])
                Fn = np.zeros((16, 4))   # [F11 F12 F21 F22] of each DOF's node
                for k, n in enumerate(nd):
                    if n >= 0:
                        Fn[2 * k] = Fn[2 * k + 1] = [F[4 * n], F[4 * n + 2], F[4 * n + 1], F[4 * n + 3]]
                be += det * (beta * np.einsum("rk,kr->k", DN, Fn) + mu * Nm.T @ Nm @ Ue)
        ok = idx >= 0
        ii, jj = np.meshgrid(idx[ok], idx[ok], indexing="ij")
        rows.append(ii.ravel()); cols.append(jj.ravel()); vals.append(Ae[np.ix_(ok, ok)].ravel())
        np.add.at(b, idx[ok], be[ok])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(2 * nn, 2 * nn))
    involved = np.unique(el[:, :8])
    has_absent = involved[0] < 0
    involved = involved[involved >= 0]
    if not has_absent:
        involved = involved[1:]      # :166 drops the first sorted entry as if it were 0
    dirichlet = np.asarray(dirichlet, np.int64).ravel()
    free_nodes = np.setdiff1d(involved, dirichlet)
    free = np.sort(np.concatenate([2 * free_nodes, 2 * free_nodes + 1]))
    Uh = np.zeros(2 * nn)
    if dirichlet.size:
        dd = np.concatenate([2 * dirichlet, 2 * dirichlet + 1])
        Uh[dd] = U[dd]
    rhs = b - A @ Uh
    if free.size:
        Uh[free] = spl.spsolve(A[free][:, free].tocsc(), rhs[free])
    return Uh


def gauss_strain_quadtree(coords, elems8, U):
    """funGlobalNodalStrainQuadtree.m:31-205 with GaussPtOrder 2: the midside nodes are
    dropped (:33, elementsFEM(:,5:8) = 0), so every element is the bilinear Q4 of its
    corners.  Gauss points x_i = g_i (x3 - x1)/2 + (x3 + x1)/2 (and y_j) from corners 1
    and 3 (:116-121), row 4 e + 2 i + j (:201-205); (ksi, eta) from the element's map
    ksi = l . [x y, x, y, 1], l = linsolve([x y, x, y, 1] of the corners, [-1 1 1 -1]')
    (eta with [-1 -1 1 1]', :102-107, :123-124); strain DN U = [u_x u_y v_x v_y]
    (:150-197).  Returns (Fgp (4E, 4), gpxy (4E, 2))."""
    coords = np.asarray(coords, np.float64).reshape(-1, 2)
    el = np.asarray(elems8, np.int64).reshape(-1, 8)[:, :4]
    U = np.asarray(U, np.float64).reshape(-1, 2)
    E = el.shape[0]
    gq = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    Fgp = np.zeros((4 * E, 4))
    gxy = np.zeros((4 * E, 2))
    for e in range(E):
        p = coords[el[e]]
        u = U[el[e]]
        L = np.column_stack([p[:, 0] * p[:, 1], p[:, 0], p[:, 1], np.ones(4)])
        lv = np.linalg.solve(L, [-1.0, 1.0, 1.0, -1.0])
        mv = np.linalg.solve(L, [-1.0, -1.0, 1.0, 1.0])
        xs = gq * 0.5 * (p[2, 0] - p[0, 0]) + 0.5 * (p[2, 0] + p[0, 0])
        ys = gq * 0.5 * (p[2, 1] - p[0, 1]) + 0.5 * (p[2, 1] + p[0, 1])
        for i in range(2):
            for j in range(2):
                x, y = xs[i], ys[j]
                ksi = lv @ [x * y, x, y, 1.0]
                eta = mv @ [x * y, x, y, 1.0]
                dk = 0.25 * np.array([-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)])
                de = 0.25 * np.array([-(1 - ksi), -(1 + ksi), 1 + ksi, 1 - ksi])
                J = np.array([[dk @ p[:, 0], dk @ p[:, 1]], [de @ p[:, 0], de @ p[:, 1]]])
                D = np.linalg.inv(J) @ np.vstack([dk, de])     # rows d/dx, d/dy
                q = 4 * e + 2 * i + j
                Fgp[q] = [D[0] @ u[:, 0], D[1] @ u[:, 0], D[0] @ u[:, 1], D[1] @ u[:, 1]]
                gxy[q] = (x, y)
    return Fgp, gxy


def gather_weights(Fgp, tri, w):
    """Node values sum_k w[i, k] Fgp[tri[i, k]] (the linear interpolation of :246-255 in
    weight form), interleaved per node."""
    tri = np.asarray(tri, np.int64).reshape(-1, 3)
    w = np.asarray(w, np.float64).reshape(-1, 3)
    return np.einsum("nk,nkc->nc", w, np.asarray(Fgp)[tri]).ravel()


def gauss_kernel(size=3, sigma=0.5):
    """fspecial('gaussian', size, sigma): exp(-(x^2+y^2)/(2 sigma^2)) over the centred
    size x size grid, entries below eps * max set to 0, normalised to sum 1."""
    r = (size - 1) / 2.0
    x = np.arange(size) - r
    h = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2.0 * sigma * sigma))
    h[h < np.finfo(float).eps * h.max()] = 0.0
    return h / h.sum()


def nanconv_edge(Z, k):
    """nanconv(Z, k, 'edge', 'nanout') (func/nanconv.m:1-...): NaNs count as 0 in the
    sum, each output is divided by the kernel weight that fell on valid in-image
    values, and NaN inputs stay NaN (conv2 'same')."""
    from scipy.signal import convolve2d
    Z = np.asarray(Z, np.float64)
    n = np.isnan(Z)
    a = np.where(n, 0.0, Z)
    flat = convolve2d((~n).astype(np.float64), k, mode="same")
    c = convolve2d(a, k, mode="same") / flat
    c[n] = np.nan
    return c


def smooth_nodal(V, coords, ncomp, passes=2, size=3, sigma=0.5):
    """funSmoothDisp.m / funSmoothStrain.m (default filter: sigma 0.5, size
    2 ceil(2 sigma) + 1 = 3): each pass fits every component by gridfit ('springs') over
    the node grid min:step:max (the unique node coordinates of a uniform mesh), blurs the
    grid with nanconv(..., 'edge', 'nanout') and reads it back at the nodes; two passes
    (SmoothTimes 1, 2)."""
    coords = np.asarray(coords, np.float64).reshape(-1, 2)
    V = np.asarray(V, np.float64).copy()
    xn, yn = np.unique(coords[:, 0]), np.unique(coords[:, 1])
    ix = np.searchsorted(xn, coords[:, 0])
    iy = np.searchsorted(yn, coords[:, 1])
    k = gauss_kernel(size, sigma)
    for _ in range(passes):
        for c in range(ncomp):
            z = V[c::ncomp]
            ok = ~np.isnan(z)
            Z = gridfit(coords[ok, 0], coords[ok, 1], z[ok], xn, yn)   # (ny, nx)
            V[c::ncomp] = nanconv_edge(Z, k)[iy, ix]
    return V


def beta_choice(beta_list, err1, err2):
    """main_ALDIC.m:285-301 (FE branch): z-scores of the two error norms (std with N-1),
    the first minimum of their sum, then the vertex of the parabola through the three
    (log10 beta, ErrSum) points around it; a minimum at either end of the list (where
    the reference's fit raises) keeps that list entry."""
    e1 = np.asarray(err1, np.float64)
    e2 = np.asarray(err2, np.float64)
    es = (e1 - e1.mean()) / e1.std(ddof=1) + (e2 - e2.mean()) / e2.std(ddof=1)
    i = int(np.argmin(es))
    if i == 0 or i == len(beta_list) - 1:
        return float(beta_list[i]), i
    x = np.log10(np.asarray(beta_list[i - 1:i + 2], np.float64))
    p = np.polyfit(x, es[i - 1:i + 2], 2)
    return float(10.0 ** (-p[1] / 2.0 / p[0])), i


# ---- whole frame (main_ALDIC.m:87-451, FEM branch, smoothing off) -------------------

def badpoint(iters, mean_rule):
    """LocalICGN.m:119-131 (mean_rule 1) / Subpb1.m:167-169 (0) / Subpb1Quadtree.m:170-181
    (2: threshold max(mean + 0.15 std, 20)): bad-subset mask."""
    it = np.asarray(iters)
    bad = (it < 0) | (it > 99)
    good = ~bad
    if mean_rule and good.any():
        m = it[good].mean()
        sd = it[good].std(ddof=1) if good.sum() > 1 else 0.0
        thr = m + 0.15 * sd
        if mean_rule == 2:
            thr = max(thr, 20.0)
        bad = bad | (it > thr)
    return bad


def sibson(pxy, q, maxn=512):
    """natural_oracle.c orc_sibson: Sibson natural-neighbour coordinates of point q with
    respect to the points pxy (n, 2), q strictly inside their convex hull -> (indices,
    weights) of the natural neighbours; raises if the cell cannot be formed."""
    pxy = _f64(pxy).reshape(-1, 2)
    idx = np.zeros(maxn, np.int32)
    w = np.zeros(maxn)
    n = lib().orc_sibson(pxy.shape[0], pxy, float(q[0]), float(q[1]), idx, w, maxn)
    if n < 0:
        raise ValueError(f"orc_sibson failed ({n})")
    return idx[:n].copy(), w[:n].copy()


def natural_fill(coords, V, nc, bad, return_kind=False):
    """LocalICGN.m:141-161 / Subpb1Quadtree.m:188-199: scatteredInterpolant(good nodes,
    'natural') of nc interleaved values per node evaluated at every node (natural_oracle.c:
    Sibson inside the good nodes' hull, the chosen linear rule on and outside it).  Good
    nodes keep their values.  kind per node: 0 good, 1 Sibson, 2 hull edge, 3 hull vertex,
    4 nearest good node (degenerate hull), 5 no good node."""
    coords = _f64(coords).reshape(-1, 2)
    nn = coords.shape[0]
    out = _f64(V).reshape(-1).copy()
    assert out.size == nc * nn
    kind = np.zeros(nn, np.int32)
    rc = lib().orc_natural_fill(nn, coords, _i32(np.asarray(bad, bool).astype(np.int32)), nc,
                                out, kind)
    if rc != 0:
        raise ValueError("orc_natural_fill: a Sibson cell could not be formed")
    return (out, kind) if return_kind else out


def frame(fraw, graw, ws, step, s, grid, tol=1e-2, mu=1e-3, beta=None, n_admm=3,
          cg_rtol=1e-10, subpb1_fit=True, smooth=False, beta_sweep=False, admm_tol=0.0,
          U0=None, mask=None, outlier_thr=100.0):
    """FP64 restatement of one frame: prep, search, LocalICGN (+ bad-point rule and
    refill), step-1 Subpb2 + strain + duals, then n_admm x (Subpb1: 2-DOF IC-GN, bad
    points, gridfit of every node; Subpb2, strain, duals).  The LocalICGN refill of bad
    nodes is the reference's scatteredInterpolant 'natural' (LocalICGN.m:144-161,
    natural_fill).  subpb1_fit=False replaces Subpb1's gridfit with the same
    natural_fill of its bad nodes.  smooth=True adds main_ALDIC.m's smoothing:
    funSmoothStrain of F1 after LocalICGN (:233) and of F2 after every global step
    (:320, :407), funSmoothDisp of U1 after Subpb1 and of U2 after Subpb2 in the ADMM
    steps (:381, :400).  beta_sweep=True picks beta as main_ALDIC.m:237,274-301 does
    (FE branch) from the step-1 global step at the six candidates.  admm_tol > 0 stops
    the ADMM steps once the RMS change of U2 or of Subpb1's U (before smoothing) between
    steps falls below it (main_ALDIC.m:409-447).  U0 (2S) replaces the integer search
    with the given initial guess (a later frame of a sequence, main_ALDIC.m:134-137);
    mask (ny x nx, 0/1) multiplies the reference gradients (funImgGradient.m:77-78).
    After the search (not with U0), outlier_thr > 0 runs IntegerSearch.m:68-76 (nodes on
    masked pixels NaN), :114 (funRemoveOutliers at Thr0 = outlier_thr, remove_outliers)
    and Init.m:13 (the remaining NaNs in-painted by springs), in those steps.
    Returns the stages' outputs and
    times, the beta used and the number of ADMM steps run."""
    import time
    import scipy.sparse as sp
    T = {}
    t0 = time.perf_counter()
    f, _ = normalize(np.asarray(fraw, np.float64))
    g, _ = normalize(np.asarray(graw, np.float64))
    gx, gy = gradient(f, mask)
    T["prep"] = time.perf_counter() - t0
    t = time.perf_counter()
    if U0 is None:
        r = search(f, g, grid, ws, step, [s, s] if np.isscalar(s) else s)
        Nx, Ny = r["Nx"], r["Ny"]
        xy = np.stack([r["x"].ravel(), r["y"].ravel()], 1)
        U0 = np.stack([r["u"].ravel(), r["v"].ravel()], 1).ravel()
        if outlier_thr > 0:
            u, v = r["u"].astype(np.float64).copy(), r["v"].astype(np.float64).copy()
            if mask is not None:
                xi, yi = r["x"], r["y"]
                integral = np.all(xi == np.floor(xi)) and np.all(yi == np.floor(yi))
                if integral:
                    hole = np.asarray(mask)[yi.astype(int) - 1, xi.astype(int) - 1] == 0
                    u[hole] = np.nan
                    v[hole] = np.nan
            u, v, _ = remove_outliers(u, v, outlier_thr)
            u, v = inpaint_springs(u), inpaint_springs(v)
            U0 = np.stack([u.ravel(), v.ravel()], 1).ravel()
    else:
        # main_ALDIC.m:134-137: a later frame starts from the previous frame's U
        ny_, nx_ = f.shape
        x1, y1, Nx, Ny = search_grid(nx_, ny_, grid, ws, step, [s, s] if np.isscalar(s) else s)
        kk, ll = np.meshgrid(np.arange(Nx), np.arange(Ny))
        xy = np.stack([(x1 + kk * step + 0.5 * ws).ravel(), (y1 + ll * step + 0.5 * ws).ravel()],
                      1).astype(np.float64)
        U0 = np.asarray(U0, np.float64).ravel().copy()
    T["search"] = time.perf_counter() - t
    S = Nx * Ny
    t = time.perf_counter()
    U6, F6, H, it6, z = icgn6(f, g, gx, gy, xy, U0, ws, tol)
    T["icgn6"] = time.perf_counter() - t
    t = time.perf_counter()
    bad = badpoint(it6, True)
    U1 = U6.copy(); F1 = F6.copy()
    U1.reshape(-1, 2)[bad] = np.nan
    F1.reshape(-1, 4)[bad] = np.nan
    U1 = natural_fill(xy, U1, 2, bad)
    F1 = natural_fill(xy, F1, 4, bad)
    T["post6"] = time.perf_counter() - t
    el = []
    for l in range(Ny - 1):
        for k in range(Nx - 1):
            n0 = l * Nx + k
            el.append([n0, n0 + 1, n0 + Nx + 1, n0 + Nx])
    el = np.array(el, np.int32)
    if beta is None:
        beta = 1e-2 * step * step * mu
    udual = np.zeros(4 * S)
    vdual = np.zeros(2 * S)
    T["subpb2"] = T["icgn2"] = 0.0
    if smooth:
        F1 = smooth_nodal(F1, xy, 4)
    if beta_sweep:
        blist = np.array([1e-3, np.sqrt(1e-5), 1e-2, np.sqrt(1e-3), 1e-1, np.sqrt(1e-1)]) \
            * step * step * mu
        e1, e2 = [], []
        for b_ in blist:
            U2k, _ = subpb2(xy, el, b_, mu, U1, F1, udual, vdual, rtol=cg_rtol)
            F2k = nodal_strain(xy, el, U2k)
            e1.append(np.linalg.norm(U1 - U2k))
            e2.append(np.linalg.norm(F1 - F2k))
        beta, _ = beta_choice(blist, e1, e2)

    def global_step(U1, F1, udual, vdual, smooth_u):
        t = time.perf_counter()
        U2, _ = subpb2(xy, el, beta, mu, U1, F1, udual, vdual, rtol=cg_rtol)
        if smooth_u:
            U2 = smooth_nodal(U2, xy, 2)
        F2 = nodal_strain(xy, el, U2)
        if smooth:
            F2 = smooth_nodal(F2, xy, 4)
        T["subpb2"] += time.perf_counter() - t
        return U2, F2, F2 - F1, U2 - U1

    U2, F2, udual, vdual = global_step(U1, F1, udual, vdual, False)
    it2 = None
    u1_prev, u2_prev = U1.copy(), U2.copy()
    upd = upd2 = np.nan
    steps = 0
    for _ in range(n_admm):
        t = time.perf_counter()
        U1, it2 = icgn2(f, g, gx, gy, xy, ws, H, mu, udual, vdual, U2, F2, tol)
        bad = badpoint(it2, False)
        U1.reshape(-1, 2)[bad] = np.nan
        if subpb1_fit:
            U1 = subpb1_smooth(U1, xy)              # Subpb1.m:185-199
        else:
            U1 = natural_fill(xy, U1, 2, bad)
        F1 = F2
        u1_saved = U1.copy()
        if smooth:
            U1 = smooth_nodal(U1, xy, 2)
        T["icgn2"] += time.perf_counter() - t
        U2, F2, udual, vdual = global_step(U1, F1, udual, vdual, smooth)
        steps += 1
        upd = np.linalg.norm(u2_prev - U2) / np.sqrt(U2.size)
        upd2 = np.linalg.norm(u1_prev - u1_saved) / np.sqrt(U2.size)
        u1_prev, u2_prev = u1_saved, U2.copy()
        if admm_tol > 0 and (upd < admm_tol or upd2 < admm_tol):
            break
    T["frame"] = time.perf_counter() - t0
    return dict(U=U2, F=F2, U6=U6, F6=F6, iters6=it6, iters2=it2, U0=U0, xy=xy, Nx=Nx, Ny=Ny,
                times=T, beta=beta, admm_steps=steps, update=(upd, upd2))


# ----------------------------------------------------------------------------------
# f4  ComputeStrain.m:11-139 on a uniform M x N node grid (node n = i + M j, x fastest)
# ----------------------------------------------------------------------------------
def derivative_op(M, N, h):
    """funDerivativeOp.m:28-67: the sparse 4MN x 2MN operator F = D U, built entry by
    entry as the reference does (border rows use the node itself as the missing
    neighbour with weights -2 / 2 instead of -1 / 1; duplicate entries summed and exact
    zeros dropped, as MATLAB's sparse() does)."""
    import scipy.sparse as sp
    I, J, V = [], [], []
    for t in range(M * N):
        x, y = t % M + 1, t // M + 1                       # ndgrid(1:M, 1:N) order
        bo = [1, 1, 1, 1]                                  # {back, front, right, left}
        if y == 1: bo[1] = 0
        if y == N: bo[0] = 0
        if x == 1: bo[3] = 0
        if x == M: bo[2] = 0
        idx = x + M * (y - 1)
        nb = [idx + bo[0] * M, idx - bo[1] * M, idx + bo[2], idx - bo[3]]
        rows = [4 * idx - 3, 4 * idx - 2, 4 * idx - 1, 4 * idx]
        c1 = [2 * nb[3] - 1, 2 * nb[3], 2 * nb[1] - 1, 2 * nb[1]]
        c2 = [2 * nb[2] - 1, 2 * nb[2], 2 * nb[0] - 1, 2 * nb[0]]
        v1 = [-1.0, -1.0, -1.0, -1.0]
        v2 = [1.0, 1.0, 1.0, 1.0]
        if bo[2] * bo[3] == 0:
            v1[0] = v1[1] = -2.0
            v2[0] = v2[1] = 2.0
        if bo[0] * bo[1] == 0:
            v1[2] = v1[3] = -2.0
            v2[2] = v2[3] = 2.0
        I += rows + rows
        J += c1 + c2
        V += v1 + v2
    A = sp.coo_matrix((V, (np.array(I) - 1, np.array(J) - 1)), shape=(4 * M * N, 2 * M * N)).tocsr()
    A.sum_duplicates()
    A.eliminate_zeros()
    return (1.0 / (2 * h)) * A


def plane_fit2(U, hx, hy, R):
    """plotFiles/PlaneFit2.m: per node (ii, jj) at least R from the border, the least
    squares of [1, ti hx, tj hy] over U(ii + ti, jj + tj), solved from its normal
    equations as the reference does ((A'A) \\ (A'b)).  U is M x N (x along rows).
    Returns (UY, UX, UNew) with the border copied from U."""
    U = np.asarray(U, np.float64)
    M, N = U.shape
    UY, UX, UNew = U.copy(), U.copy(), U.copy()
    L = np.ones(((2 * R + 1) ** 2, 3))
    for tj in range(-R, R + 1):
        for ti in range(-R, R + 1):
            L[(2 * R + 1) * (tj + R) + ti + R] = [1.0, ti * hx, tj * hy]
    AtA = L.T @ L
    for ii in range(R, M - R):
        for jj in range(R, N - R):
            b = np.empty((2 * R + 1) ** 2)
            for tj in range(-R, R + 1):
                for ti in range(-R, R + 1):
                    b[(2 * R + 1) * (tj + R) + ti + R] = U[ii + ti, jj + tj]
            v = np.linalg.solve(AtA, L.T @ b)
            UNew[ii, jj], UX[ii, jj], UY[ii, jj] = v
    return UY, UX, UNew


def strain_type(Fs, t):
    """ComputeStrain.m:98-136: infinitesimal (0), Eulerian (1), Green-Lagrangian (2)."""
    F = np.array(Fs, np.float64)
    ux, vx, uy, vy = (np.asarray(Fs, np.float64)[k::4] for k in range(4))
    if t == 1:
        F[0::4] = 1.0 / (1.0 - ux) - 1.0
        F[3::4] = 1.0 / (1.0 - vy) - 1.0
        F[2::4] = uy / (1.0 - vy)
        F[1::4] = vx / (1.0 - ux)
    elif t == 2:
        F[0::4] = 0.5 * (ux * 2 - ux ** 2 - vx ** 2)
        F[3::4] = 0.5 * (vy * 2 - uy ** 2 - vy ** 2)
        F[2::4] = 0.5 * (uy + vx - ux * uy - vx * vy)
        F[1::4] = 0.5 * (vx + uy - uy * ux - vy * vx)
    elif t != 0:
        raise ValueError("strain type 0, 1 or 2")
    return F


def trim_index(M, N, r):
    """ComputeStrain.m:18-24: temp3, the 0-based F entries of the nodes at least r from
    the border, reshape(1:MN, M, N)(r+1:M-r, r+1:N-r) in column-major order."""
    temp = np.arange(M * N).reshape(N, M).T               # reshape(temp, M, N), 0-based
    t2 = temp[r:M - r, r:N - r].flatten(order="F")
    return (4 * t2[:, None] + np.arange(4)[None, :]).ravel()


def compute_strain(M, N, h, U=None, F=None, method=1, rad=1, stype=0, coords=None, elems=None):
    """ComputeStrain.m:11-139: FStrain by the method, the strain type, then the trimmed
    FStraintemp and FStrainWorld (F21 and F12 negated).  Method 3 uses nodal_strain
    (funGlobal_NodalStrainAvg) on the given mesh.  Returns a dict."""
    if method == 0:
        Fs, r = np.array(F, np.float64), 0
    elif method == 1:
        Fs, r = derivative_op(M, N, h) @ np.asarray(U, np.float64), 1
    elif method == 2:
        U = np.asarray(U, np.float64)
        Fs, r = derivative_op(M, N, h) @ U, rad
        uy, ux, _ = plane_fit2(U[0::2].reshape(N, M).T, h, h, rad)
        vy, vx, _ = plane_fit2(U[1::2].reshape(N, M).T, h, h, rad)
        n = (M - 2 * rad) * (N - 2 * rad)
        ft = np.zeros(4 * n)
        ft[0::4] = ux[rad:M - rad, rad:N - rad].flatten(order="F")
        ft[1::4] = vx[rad:M - rad, rad:N - rad].flatten(order="F")
        ft[2::4] = uy[rad:M - rad, rad:N - rad].flatten(order="F")
        ft[3::4] = vy[rad:M - rad, rad:N - rad].flatten(order="F")
        Fs[trim_index(M, N, rad)] = ft
    elif method == 3:
        Fs, r = nodal_strain(coords, elems, U), 1
    else:
        raise ValueError("method 0..3")
    Ff = strain_type(Fs, stype)
    t3 = trim_index(M, N, r)
    Ft = Ff[t3]
    Fw = Ft.copy()
    Fw[1::4] = -Fw[1::4]
    Fw[2::4] = -Fw[2::4]
    return dict(FStrain=Fs, FStrainFinite=Ff, FStraintemp=Ft, FStrainWorld=Fw, Rad=r, temp3=t3)


# ----------------------------------------------------------------------------------
# f2  funRemoveOutliers.m:45-86,371-445 (non-interactive median test) and
#     inpaint_nans.m methods 3 and 4 (Init.m:13)
# ----------------------------------------------------------------------------------
def inpaint_springs(A):
    """inpaint_nans.m method 4: a spring between each NaN element and each of its in-array
    horizontal and vertical neighbours (each unique pair once), the known values moved to
    the right-hand side, the NaNs solved in least squares."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    A = np.array(A, dtype=np.float64)
    nan = np.isnan(A)
    if not nan.any():
        return A
    n, m = A.shape
    idx = np.arange(n * m).reshape(n, m)
    pairs = set()
    for (i, j) in zip(*np.nonzero(nan)):
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            a, b = i + di, j + dj
            if 0 <= a < n and 0 <= b < m:
                pairs.add(tuple(sorted((idx[i, j], idx[a, b]))))
    rows, cols, vals = [], [], []
    for k, (p, q) in enumerate(sorted(pairs)):
        rows += [k, k]
        cols += [p, q]
        vals += [1.0, -1.0]
    S = sp.csr_matrix((vals, (rows, cols)), shape=(len(pairs), n * m))
    flat = A.ravel()
    known = ~nan.ravel()
    rhs = -S[:, known] @ flat[known]
    unk = np.nonzero(~known)[0]
    Su = S[:, unk]
    sol = spl.spsolve((Su.T @ Su).tocsc(), Su.T @ rhs)
    flat = flat.copy()
    flat[unk] = sol
    return flat.reshape(n, m)


def inpaint_plate(A):
    """inpaint_nans.m method 3: the del^4 template on every element the NaNs' 13-point
    template reaches, reduced near the border (the 5-point Laplacian on the second ring,
    second differences along the first row and column, nothing at the corners), the
    known values moved to the right-hand side and the NaNs solved in least squares."""
    A = np.array(A, dtype=np.float64)
    nan = np.isnan(A)
    if not nan.any():
        return A
    n, m = A.shape
    talks = [(-2, 0), (-1, -1), (-1, 0), (-1, 1), (0, -2), (0, -1), (0, 1), (0, 2), (1, -1),
             (1, 0), (1, 1), (2, 0)]
    nodes = set(zip(*np.nonzero(nan)))
    for (i, j) in list(nodes):
        for di, dj in talks:
            if 0 <= i + di < n and 0 <= j + dj < m:
                nodes.add((i + di, j + dj))
    d4 = [((0, -2), 1), ((-1, -1), 2), ((0, -1), -8), ((1, -1), 2), ((-2, 0), 1), ((-1, 0), -8),
          ((0, 0), 20), ((1, 0), -8), ((2, 0), 1), ((-1, 1), 2), ((0, 1), -8), ((1, 1), 2),
          ((0, 2), 1)]
    lap = [((0, -1), 1), ((-1, 0), 1), ((0, 0), -4), ((1, 0), 1), ((0, 1), 1)]
    rows = []
    for (i, j) in sorted(nodes):
        r, c = i + 1, j + 1
        if 3 <= r <= n - 2 and 3 <= c <= m - 2:
            st = d4
        elif ((r in (2, n - 1)) and 2 <= c <= m - 1) or ((c in (2, m - 1)) and 2 <= r <= n - 1):
            st = lap
        elif r in (1, n) and 2 <= c <= m - 1:
            st = [((0, -1), 1), ((0, 0), -2), ((0, 1), 1)]
        elif c in (1, m) and 2 <= r <= n - 1:
            st = [((-1, 0), 1), ((0, 0), -2), ((1, 0), 1)]
        else:
            continue
        rows.append([((i + di, j + dj), w) for (di, dj), w in st])
    unk = sorted(zip(*np.nonzero(nan)))
    col = {p: k for k, p in enumerate(unk)}
    M = np.zeros((len(rows), len(unk)))
    rhs = np.zeros(len(rows))
    for k, st in enumerate(rows):
        for p, w in st:
            if p in col:
                M[k, col[p]] += w
            else:
                rhs[k] -= w * A[p]
    keep = np.any(M != 0.0, axis=1)
    sol = np.linalg.lstsq(M[keep], rhs[keep], rcond=None)[0]
    for p, k in col.items():
        A[p] = sol[k]
    return A


def medfilt3(V):
    """medFilt2(V, [2 2], 'symmetric', 2) as funRemoveOutliers.m:373-376 calls it: one
    mirrored row and column after the array, the 2 x 2 neighbourhood without its second
    member, i.e. the median of V(i, j), V(i+1, j), V(i+1, j+1), in single precision
    (the neighbour stack is single, :404)."""
    P = np.pad(np.asarray(V), ((0, 1), (0, 1)), mode="symmetric").astype(np.float32)
    st = np.sort(np.stack([P[:-1, :-1], P[1:, :-1], P[1:, 1:]]), axis=0)
    return st[1]


def med_fluct(u, epsilon=0.1):
    """funRemoveOutliers.m:371-379 (funMedRemoveOutliers): NaNs in-painted by method 3,
    then the normalised fluctuation |fluct ./ (medianRes + epsilon)|.  The medians are
    single, so (MATLAB's mixed arithmetic) fluct and everything after it are single."""
    u = inpaint_plate(u)
    med = medfilt3(u)
    fluct = u.astype(np.float32) - med
    res = medfilt3(np.abs(fluct))
    nf = np.abs(fluct / (res + np.float32(epsilon)))
    return med, nf


def remove_outliers(u, v, thr=100.0):
    """funRemoveOutliers(u, v, cc, 0.5, Thr0) without its prompts (:45-86): the median
    test of u and v, elements with sqrt(nf_u^2 + nf_v^2) > Thr0 set NaN and in-painted
    by springs together with the NaNs, then the NaNs put back.  Returns (u, v, flagged)."""
    u = np.array(u, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    hole = np.isnan(u)
    _, nfu = med_fluct(u)
    _, nfv = med_fluct(v)
    mag = np.sqrt(nfu * nfu + nfv * nfv)                   # single, :50-51
    flagged = mag > np.float32(thr)
    u2, v2 = u.copy(), v.copy()
    u2[flagged] = np.nan
    v2[flagged] = np.nan
    u2 = inpaint_springs(u2)
    v2 = inpaint_springs(v2)
    u2[hole] = np.nan
    v2[hole] = np.nan
    return u2, v2, flagged
