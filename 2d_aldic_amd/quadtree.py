"""Host-side mirror of the reference's quadtree ADMM steps (func_quadtree/*.m).

Same names and argument meaning as the reference functions; the compute runs in the HIP
library (include/aldic.h: aldic_set_quadtree, aldic_subpb2_quadtree,
aldic_nodal_strain_quadtree, aldic_icgn2).  This module holds the host geometry the
reference leaves to MATLAB's scatteredInterpolant 'linear' (the nodal strain's Gauss-point
interpolation weights of a Delaunay triangulation, built once per mesh, like the mesh
itself); its 'natural' refill of bad nodes runs in the library (aldic_natural_fill).

Mesh convention: DICmesh['coordinatesFEM'] (nn, 2); DICmesh['elementsFEM'] (E, 8),
1-based, 0 where an edge has no midside node (GenerateQuadtreeMesh.m); optional
DICmesh['dirichlet'] (1-based node list).
"""
from __future__ import annotations

import time

import numpy as np

from .aldic import _ensure_images, _fingerprint, get_context

_QT_KEY = None
_W_CACHE = {}


def linear_weights(points, queries):
    """Weights of piecewise-linear interpolation over the Delaunay triangulation of
    `points` at `queries`: (tri (nq, 3) point indices, w (nq, 3)), sum_k w = 1.  A query
    outside the convex hull takes the plane of the triangle on the nearest hull edge
    (linear extrapolation; the first such edge on ties)."""
    from scipy.spatial import Delaunay
    P = np.asarray(points, np.float64).reshape(-1, 2)
    Q = np.asarray(queries, np.float64).reshape(-1, 2)
    dt = Delaunay(P)
    s = dt.find_simplex(Q)
    out = s < 0
    if out.any():
        faces = []  # hull edges with their triangle
        for t, nb in enumerate(dt.neighbors):
            for k in range(3):
                if nb[k] == -1:
                    a, b = np.delete(dt.simplices[t], k)
                    faces.append((a, b, t))
        fa = np.array(faces)
        A, B = P[fa[:, 0]], P[fa[:, 1]]
        AB = B - A
        L2 = np.einsum("ij,ij->i", AB, AB)
        for qi in np.flatnonzero(out):
            AQ = Q[qi] - A
            t = np.clip(np.einsum("ij,ij->i", AQ, AB) / L2, 0.0, 1.0)
            d = AQ - t[:, None] * AB
            s[qi] = fa[int(np.argmin(np.einsum("ij,ij->i", d, d))), 2]
    T = dt.transform[s]
    b = np.einsum("kij,kj->ki", T[:, :2], Q - T[:, 2])
    w = np.column_stack([b, 1.0 - b.sum(axis=1)])
    return dt.simplices[s].astype(np.int32), w


def scattered_fill(coords, V, nc, bad):
    """scatteredInterpolant(good nodes, 'natural') at the bad nodes (Subpb1Quadtree.m:
    188-199) on the device (aldic_natural_fill); nc = 2 (U) or 4 (F) values per node."""
    coords = np.asarray(coords, np.float64)
    bad = np.asarray(bad, bool)
    V = np.asarray(V, np.float64).ravel()
    if not bad.any():
        return V.copy()
    ctx = get_context()
    if nc == 2:
        return ctx.natural_fill(coords, bad, V)[0]
    if nc == 4:
        U0 = np.zeros(2 * coords.shape[0])
        return ctx.natural_fill(coords, bad, U0, V)[1]
    raise ValueError("nc must be 2 or 4")


def _ensure_quadtree(DICmesh):
    global _QT_KEY
    ctx = get_context()
    coords = np.asarray(DICmesh["coordinatesFEM"], np.float64)
    el = np.asarray(DICmesh["elementsFEM"], np.int64)
    d = np.asarray(DICmesh.get("dirichlet", []), np.int64).ravel()
    key = _fingerprint(coords, el, d)
    if key != _QT_KEY:
        el8 = np.full((el.shape[0], 8), -1, np.int64)
        el8[:, :el.shape[1]] = el - 1          # 1-based, 0 absent -> 0-based, -1 absent
        ctx.set_quadtree(coords, el8.astype(np.int32), (d - 1).astype(np.int32))
        _QT_KEY = key
    return ctx, key


def Subpb2Quadtree(DICmesh, GaussPtOrder, beta, mu, U, F, udual, vdual, alpha=0.0, hbar=None,
                   ClusterNo=1, rtol=1e-12):
    """Subpb2Quadtree.m:1-346 -> Uhat.  The duals enter multiplied by 0 (:42-43);
    main_ALDIC_Quadtree.m:181 sets GaussPtOrder 2 and alpha 0, the only form provided.
    The sparse direct solve becomes FP64 Jacobi-PCG to relative residual rtol."""
    if GaussPtOrder != 2 or alpha != 0:
        raise NotImplementedError("Subpb2Quadtree is provided for GaussPtOrder 2, alpha 0")
    ctx, _ = _ensure_quadtree(DICmesh)
    Uh, _ = ctx.subpb2_quadtree(beta, mu, U, F, rtol=rtol)
    return Uh


def funGlobalNodalStrainQuadtree(DICmesh, U, GaussPtOrder=2, waitBarDisplayOrNot=None):
    """funGlobalNodalStrainQuadtree.m:1-296 -> [StrainNodalPt, StrainGaussPt,
    CoordGaussPt] (smoothness 0: Gauss-point strains, then linear interpolation at the
    nodes with the weights of linear_weights)."""
    if GaussPtOrder != 2:
        raise NotImplementedError("funGlobalNodalStrainQuadtree is provided for GaussPtOrder 2")
    ctx, key = _ensure_quadtree(DICmesh)
    if key not in _W_CACHE:
        _, _, gxy = ctx.nodal_strain_quadtree(U)
        _W_CACHE.clear()
        _W_CACHE[key] = linear_weights(gxy, DICmesh["coordinatesFEM"])
    tri, w = _W_CACHE[key]
    F, Fgp, gxy = ctx.nodal_strain_quadtree(U, tri, w)
    return F, Fgp, gxy


def Subpb1Quadtree(UOld, FOld, udual, vdual, coordinatesFEM, Df, ImgRef, ImgDef, mu, beta, HPar,
                   ALSolveStep, DICpara, ICGNmethod="GaussNewton", tol=1e-2):
    """Subpb1Quadtree.m:1-199 -> [U, HPar, ALSub1Time, ConvItPerEle, LocalICGNBadPtNum].
    Window list mean(winsize) at every node (:50); bad = status < 0 or > 99, then also
    status > max(mean + 0.15 std, 20) over the good ones (:170-181); bad nodes refilled
    by scattered_fill."""
    ctx = _ensure_images(ImgRef, ImgDef, Df)
    coords = np.asarray(coordinatesFEM, np.float64)
    S = coords.shape[0]
    H = np.asarray(HPar, np.float64)
    if H.shape == (21, S):
        H = H.T
    if ALSolveStep <= 1:
        H = np.zeros((S, 21))
    ws = int(round(float(np.mean(DICpara["winsize"]))))
    t0 = time.perf_counter()
    U, it = ctx.icgn2(coords, ws, H, mu, vdual, UOld, FOld, tol)
    ALSub1Time = time.perf_counter() - t0
    it = it.astype(np.int64)
    bad = (it < 0) | (it > 99)
    LocalICGNBadPtNum = int(bad.sum()) - int(np.sum(it == 102))
    good = ~bad
    if good.any():
        m = it[good].mean()
        sd = it[good].std(ddof=1) if good.sum() > 1 else 0.0
        bad = bad | (it > max(m + 0.15 * sd, 20.0))
    U = U.copy()
    U[0::2][bad] = np.nan
    U[1::2][bad] = np.nan
    U = scattered_fill(coords, U, 2, bad)
    return U, HPar, ALSub1Time, it, LocalICGNBadPtNum


__all__ = ["linear_weights", "scattered_fill", "Subpb2Quadtree", "funGlobalNodalStrainQuadtree",
           "Subpb1Quadtree"]
