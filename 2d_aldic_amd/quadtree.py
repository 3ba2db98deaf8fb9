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


# ---- quadtree mesh generation (host geometry, like MeshSetUp) ---------------------------

def _mask_boxes(coordinates, elements, ImgRefMask):
    """Per element: the 1-based pixel box [xmin, xmax] x [ymin, ymax] spanned by its four
    corners, and the sum and the min/max of ImgRefMask (MATLAB shape, dim 1 = x) over the
    box, inclusive, as funMarkEdge.m and funMarkInside.m index it.  Sums and a constant
    test come from an integral image."""
    c = np.asarray(coordinates, np.float64)
    el = np.asarray(elements, np.int64)[:, :4] - 1
    X, Y = c[el, 0], c[el, 1]
    xmn, xmx, ymn, ymx = X.min(1), X.max(1), Y.min(1), Y.max(1)
    m = np.asarray(ImgRefMask, np.float64)
    for v in (xmn, xmx, ymn, ymx):
        if np.any(v != np.round(v)):
            raise ValueError("quadtree mesh: element corners must sit on whole pixels")
    x0, x1 = xmn.astype(np.int64) - 1, xmx.astype(np.int64)
    y0, y1 = ymn.astype(np.int64) - 1, ymx.astype(np.int64)
    if np.any(x0 < 0) or np.any(y0 < 0) or np.any(x1 > m.shape[0]) or np.any(y1 > m.shape[1]):
        raise ValueError("quadtree mesh: an element reaches outside the mask")
    I = np.zeros((m.shape[0] + 1, m.shape[1] + 1))
    I[1:, 1:] = m.cumsum(0).cumsum(1)
    ssum = I[x1, y1] - I[x0, y1] - I[x1, y0] + I[x0, y0]
    lo, hi = m.min(), m.max()
    I2 = np.zeros_like(I)
    I2[1:, 1:] = (m == lo).cumsum(0).cumsum(1)
    nlo = I2[x1, y1] - I2[x0, y1] - I2[x1, y0] + I2[x0, y0]
    npx = (x1 - x0) * (y1 - y0)
    if np.all((m == lo) | (m == hi)):          # two-valued mask: range from the counts
        rng = np.where((nlo > 0) & (nlo < npx), hi - lo, 0.0)
    else:
        rng = np.array([np.ptp(m[a:b, p:q]) for a, b, p, q in zip(x0, x1, y0, y1)])
    return (xmn, xmx, ymn, ymx), ssum, rng


def funMarkEdge(coordinates, elements4, ImgRefMask, elementMinSize):
    """funMarkEdge.m:1-50 -> [mark3, mark4]: an element is marked when ImgRefMask is not
    constant over its pixel box and min(width, height) > elementMinSize.  mark3 = []
    (triangles are not coded in the reference either); mark4 is a boolean per element."""
    el = np.asarray(elements4)
    if el.size == 0:
        return [], np.zeros(0, bool)
    (xmn, xmx, ymn, ymx), _, rng = _mask_boxes(coordinates, el, ImgRefMask)
    size = np.minimum(np.abs(xmx - xmn), np.abs(ymx - ymn))
    return [], (rng > 0) & (size > elementMinSize)


def funMarkInside(coordinates, elements4, ImgRefMask):
    """funMarkInside.m:1-38 -> [markInside4, markOutside4] (1-based): inside when the mask
    sum over the element's pixel box is below half of (xmax - xmin)(ymax - ymin)."""
    el = np.asarray(elements4)
    if el.size == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    (xmn, xmx, ymn, ymx), ssum, _ = _mask_boxes(coordinates, el, ImgRefMask)
    inside = ssum < 0.5 * (xmx - xmn) * (ymx - ymn)
    return np.flatnonzero(inside) + 1, np.flatnonzero(~inside) + 1


def _column_major_nonzero(A):
    """(rows, cols) of the nonzero entries of A in column-major order (MATLAB's find)."""
    c, r = np.nonzero(np.asarray(A).T)
    return r, c


def provideGeometricData(irregular, elements):
    """provideGeometricData.m for (elements3 = irregular, elements4): the edges of each
    irregular triple (n1, n2), (n2, n3), (n3, n1) and of each quadrilateral (1, 2), (2, 3),
    (3, 4), (4, 1), numbered by the sorted unique rows of the node pairs (each pair sorted
    ascending).  Returns edge2nodes (1-based nodes), irregular2edges, element2edges
    (0-based edge numbers)."""
    ir = np.asarray(irregular, np.int64).reshape(-1, 3)
    el = np.asarray(elements, np.int64)[:, :4]
    E = np.concatenate([ir[:, [0, 1]], ir[:, [1, 2]], ir[:, [2, 0]],
                        el[:, [0, 1]], el[:, [1, 2]], el[:, [2, 3]], el[:, [3, 0]]])
    E.sort(axis=1)
    edge2nodes, ie = np.unique(E, axis=0, return_inverse=True)
    ie = np.asarray(ie).ravel()
    ni = ir.shape[0]
    irregular2edges = ie[:3 * ni].reshape(3, ni).T
    element2edges = ie[3 * ni:].reshape(4, el.shape[0]).T
    return edge2nodes, irregular2edges, element2edges


def QrefineR(coordinates, elements, irregular, marked):
    """QrefineR.m (red refinement with hanging nodes, at most one per edge) -> [coordinates,
    newElements, newIrregular].  Marked elements (1-based) are split in four through the
    edge midpoints and the centroid; the closure refines every element with three marked
    edges, or whose hanging-node edge would receive a second level, until nothing
    changes.  New edge nodes are numbered in edge order, then the centroids in element
    order; the children of element e = (1, 2, 3, 4) with edge nodes m1..m4 and centroid c
    are (1, m1, c, m4), (2, m2, c, m1), (3, m3, c, m2), (4, m4, c, m3) in e's place.
    irregular rows are (edge end, edge end, hanging node), 1-based."""
    coords = np.asarray(coordinates, np.float64).reshape(-1, 2)
    el = np.asarray(elements, np.int64)[:, :4]
    irr = np.asarray(irregular, np.int64).reshape(-1, 3)
    marked = np.asarray(marked, np.int64).ravel()
    nE = el.shape[0]
    e2nodes, irr2ed, el2ed = provideGeometricData(irr, el)
    e2n = np.zeros(e2nodes.shape[0], np.int64)
    e2n[el2ed[marked - 1].ravel()] = 1
    e2n[irr2ed[:, 0]] = 1
    while True:
        me = e2n[el2ed]
        s = np.abs(me).sum(axis=1)
        kdx = np.flatnonzero((s < 4) & ((s > 2) | (me.min(axis=1, initial=0) < 0)))
        if kdx.size:
            e2n[el2ed[kdx][me[kdx] == 0]] = 1
        me2 = e2n[irr2ed]
        flag = irr2ed[np.any(me2[:, 1:] != 0, axis=1), 0]
        swap = np.flatnonzero(e2n[flag] != -1)
        e2n[flag[swap]] = -1
        if kdx.size == 0 and swap.size == 0:
            break
    e2n[irr2ed[:, 0]] = -1
    idx = e2n > 0
    nc = coords.shape[0]
    e2n[idx] = nc + np.arange(1, int(idx.sum()) + 1)
    coords = np.vstack([coords, 0.5 * (coords[e2nodes[idx, 0] - 1] + coords[e2nodes[idx, 1] - 1])])
    e2n[irr2ed[:, 0]] = irr[:, 2]
    nn = e2n[el2ed]
    reftyp = (nn != 0).astype(np.int64) @ np.array([1, 2, 4, 8])
    none, red = reftyp < 15, reftyp == 15
    ridx = np.flatnonzero(red)
    mid = np.zeros(nE, np.int64)
    mid[ridx] = coords.shape[0] + np.arange(1, ridx.size + 1)
    coords = np.vstack([coords, coords[el[ridx] - 1].mean(axis=1)])
    cnt = np.where(red, 4, 1)
    start = np.concatenate([[0], np.cumsum(cnt)])
    newEl = np.zeros((int(start[-1]), 4), np.int64)
    newEl[start[:-1][none]] = el[none]
    r, st = ridx, start[:-1][ridx]
    for k in range(4):
        newEl[st + k] = np.column_stack([el[r, k], nn[r, k], mid[r], nn[r, (k + 3) % 4]])
    kdx = np.flatnonzero((reftyp > 0) & (reftyp < 15))
    ii, jj = _column_major_nonzero(nn[kdx])
    edx = el2ed[kdx[ii], jj]
    part1 = np.column_stack([e2nodes[edx], nn[kdx][ii, jj]])
    nn2 = e2n[irr2ed[:, 1:3]]
    kdx = np.flatnonzero(nn2.sum(axis=1) != 0)
    ii, jj = _column_major_nonzero(nn2[kdx])
    edx = irr2ed[kdx[ii], jj + 1]
    part2 = np.column_stack([e2nodes[edx], nn2[kdx][ii, jj]])
    newIrregular = np.vstack([part1.reshape(-1, 3), part2.reshape(-1, 3)]).astype(np.int64)
    return coords, newEl, newIrregular


def _hanging_columns(elements4, irregular):
    """GenerateQuadtreeMesh.m:39-56: each hanging node (a, b, h) goes into column 8, 5, 6
    or 7 (1-based) of the element whose edge (1,2), (2,3), (3,4) or (4,1) joins a and b, in
    either direction (the first such element, as ismember(..., 'rows') returns)."""
    el = np.asarray(elements4, np.int64)[:, :4]
    out = np.zeros((el.shape[0], 8), np.int64)
    out[:, :4] = el
    first = {}
    for col, (p, q) in ((7, (0, 1)), (4, (1, 2)), (5, (2, 3)), (6, (3, 0))):
        for e in range(el.shape[0] - 1, -1, -1):     # lowest element index wins
            first[(col, el[e, p], el[e, q])] = e
            first[(col, -el[e, q], -el[e, p])] = e   # reversed direction, own key space
    for a, b, h in np.asarray(irregular, np.int64).reshape(-1, 3):
        for col in (7, 4, 5, 6):
            for key in ((col, a, b), (col, -a, -b)):
                e = first.get(key)
                if e is not None:
                    out[e, col] = h
    return out


def GenerateQuadtreeMesh(DICmesh, Df, DICpara, U0=None):
    """GenerateQuadtreeMesh.m:11-150 without its figures and the interactive
    funRemoveOutliersQuadtree that ends it -> (DICmesh, U0).

    Starting from the uniform DICmesh (MeshSetUp), refine by QrefineR every element whose
    pixel box crosses the mask boundary and is wider than 2 * elementMinSize (default 2,
    main_ALDIC_Quadtree.m:32), until none is marked; put the hanging nodes into columns 5-8;
    drop the elements inside holes (funMarkInside); mark the refined elements (diagonal
    below 0.99 sqrt(2) winstepsize) and those within 1.01 winstepsize of the uniform mesh's
    border, take their nodes, and grow that set twice by element neighbours: that is
    markCoordHoleEdge and dirichlet.  Nodes inside holes stay in coordinatesFEM (in no
    element), as in the reference.  U0 (interleaved, on the uniform nodes) is carried to the
    new nodes by scatteredInterpolant's piecewise-linear rule (linear_weights); a node at a
    cell centre lies on both diagonals of four cocircular nodes, where the value depends on
    the triangulation's choice of diagonal.  Elements and node lists are 1-based."""
    mask = np.asarray(Df["ImgRefMask"] if isinstance(Df, dict) else Df, np.float64)
    step = float(np.max(np.atleast_1d(DICpara["winstepsize"])))
    coords0 = np.asarray(DICmesh["coordinatesFEM"], np.float64)
    minsize = DICmesh.get("elementMinSize", 2)
    coordsQ = coords0.copy()
    elQ = np.asarray(DICmesh["elementsFEM"], np.int64)[:, :4]
    irregular = np.zeros((0, 3), np.int64)
    while True:
        _, mark4 = funMarkEdge(coordsQ, elQ, mask, minsize * 2)
        mark4 = np.flatnonzero(mark4) + 1
        coordsQ, elQ, irregular = QrefineR(coordsQ, elQ, irregular, mark4)
        if mark4.size == 0:
            break
    el8 = _hanging_columns(elQ, irregular)
    _, outside = funMarkInside(coordsQ, el8, mask)
    el8 = el8[outside - 1]
    c1, c3 = coordsQ[el8[:, 0] - 1], coordsQ[el8[:, 2] - 1]
    size = np.sqrt(((c1 - c3) ** 2).sum(axis=1))
    xmn, xmx = coords0[:, 0].min(), coords0[:, 0].max()
    ymn, ymx = coords0[:, 1].min(), coords0[:, 1].max()
    edge = ((size < 0.99 * np.sqrt(2) * step) | (c1[:, 0] < xmn + 1.01 * step)
            | (c3[:, 0] > xmx - 1.01 * step) | (c1[:, 1] < ymn + 1.01 * step)
            | (c3[:, 1] > ymx - 1.01 * step))
    hole_edge = np.setdiff1d(np.unique(el8[edge]), [0])
    for _ in range(2):
        touch = np.isin(el8, hole_edge).any(axis=1)
        hole_edge = np.setdiff1d(np.unique(el8[touch]), [0])
    out = dict(DICmesh)
    out.update(coordinatesFEM=coordsQ, elementsFEM=el8, markCoordHoleEdge=hole_edge,
               dirichlet=hole_edge.copy(), elementMinSize=minsize,
               coordinatesFEMWorld=np.column_stack([coordsQ[:, 0],
                                                    mask.shape[1] + 1 - coordsQ[:, 1]]))
    if U0 is None:
        return out, None
    U0 = np.asarray(U0, np.float64).ravel()
    tri, w = linear_weights(coords0, coordsQ)
    Un = np.empty(2 * coordsQ.shape[0])
    Un[0::2] = (U0[0::2][tri] * w).sum(axis=1)
    Un[1::2] = (U0[1::2][tri] * w).sum(axis=1)
    return out, Un


__all__ = ["linear_weights", "scattered_fill", "Subpb2Quadtree", "funGlobalNodalStrainQuadtree",
           "Subpb1Quadtree", "funMarkEdge", "funMarkInside", "provideGeometricData", "QrefineR",
           "GenerateQuadtreeMesh"]
