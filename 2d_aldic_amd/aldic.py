"""Host-side mirror of the reference's MATLAB call surface for the correlation path.

Each function keeps the name, argument order and returned values of the reference
function it replaces (jyang526843/2D_ALDIC, cited per function), so a port of
main_ALDIC.m reads the same.  Compute runs in the gfx950 HIP library through the C ABI
(include/aldic.h); this module only reshapes arrays and applies the host-side glue
(mesh set-up, bad-point bookkeeping).  There is no CPU fallback: without the built
library and a GPU every compute call raises.

Array conventions are MATLAB's after the reference's image transpose: an image is an
(nx, ny) array indexed Img[x-1, y-1]; grids such as x0, u, v are (Ny, Nx) like the
reference's; nodal vectors are interleaved (U = [u1 v1 ...], F = [F11 F21 F12 F22 ...]);
HPar / HtempPar hold the 21 lower-triangle Hessian entries per node.
"""
from __future__ import annotations

import time

import numpy as np

from ._lib import AldicError, Context

_CTX = None
_IMG_KEY = None
_GRAD_KEY = None


def get_context(device=0) -> Context:
    global _CTX
    if _CTX is None:
        _CTX = Context(device)
    return _CTX


def _cimg(Img):
    """MATLAB-shaped (nx, ny) image -> C-ordered (ny, nx) array (x fastest)."""
    return np.ascontiguousarray(np.asarray(Img).T)


def _fingerprint(*arrs):
    key = []
    for a in arrs:
        a = np.asarray(a)
        flat = a.ravel(order="K")
        step = max(1, flat.size // 4096)
        key.append((a.shape, a.dtype.str, a.__array_interface__["data"][0],
                    float(np.nansum(flat[::step].astype(np.float64)))))
    return tuple(key)


def _ensure_images(ImgRef, ImgDef, Df=None):
    """Upload the normalised pair (and Df's gradients) unless already resident."""
    global _IMG_KEY, _GRAD_KEY
    ctx = get_context()
    key = _fingerprint(ImgRef, ImgDef)
    if key != _IMG_KEY:
        ctx.set_normalized(_cimg(ImgRef), _cimg(ImgDef))
        _IMG_KEY = key
        _GRAD_KEY = None
    if Df is not None:
        gkey = _fingerprint(Df["DfDx"], Df["DfDy"]) + key
        if gkey != _GRAD_KEY:
            nx, ny = np.asarray(ImgRef).shape
            gx = np.zeros((nx, ny))
            gy = np.zeros((nx, ny))
            c = int(Df.get("DfCropWidth", 3))
            gx[c:nx - c, c:ny - c] = Df["DfDx"]
            gy[c:nx - c, c:ny - c] = Df["DfDy"]
            ctx.set_gradients(_cimg(gx), _cimg(gy))
            _GRAD_KEY = gkey
    return ctx


# ---- pre-processing ------------------------------------------------------------------

def funNormalizeImg(Img, gridxyROIRange):
    """funNormalizeImg.m:1-40 -- each image (I - mean_ROI) / std_ROI (std N-1)."""
    ctx = get_context()
    gx = list(gridxyROIRange["gridx"])
    gy = list(gridxyROIRange["gridy"])
    nx, ny = np.asarray(Img[0]).shape
    gx = [max(1, int(gx[0])), min(nx, int(gx[1]))]
    gy = [max(1, int(gy[0])), min(ny, int(gy[1]))]
    out = []
    for im in Img:
        a = _cimg(np.asarray(im, dtype=np.float64))
        ctx.set_images(a, a, roi=[gx[0], gx[1], gy[0], gy[1]])
        f, _, _, _ = ctx.get_images()
        out.append(np.ascontiguousarray(f.T))
    global _IMG_KEY
    _IMG_KEY = None
    return out, dict(gridx=gx, gridy=gy)


def funImgGradient(ImgRef, ImgDef, ImgRefMask=None):
    """funImgGradient.m:53-78 -- 7-tap gradients of ImgRef, cropped by 3, times the mask."""
    ctx = _ensure_images(ImgRef, ImgDef)
    nx, ny = np.asarray(ImgRef).shape
    _, _, gx, gy = ctx.get_images()
    gx = gx.T
    gy = gy.T
    if ImgRefMask is None:
        ImgRefMask = np.ones((nx, ny))
    m = np.asarray(ImgRefMask, dtype=np.float64)
    Df = dict(DfCropWidth=3, DfAxis=[4, nx - 3, 4, ny - 3], imgSize=(nx, ny),
              ImgRefMask=ImgRefMask,
              DfDx=gx[3:nx - 3, 3:ny - 3] * m[3:nx - 3, 3:ny - 3],
              DfDy=gy[3:nx - 3, 3:ny - 3] * m[3:nx - 3, 3:ny - 3])
    return Df


def ba_interp2(F, X, Y, method="cubic"):
    """ba_interp2.cpp:318-335 -- Keys cubic, clamp-to-edge.  X indexes dim 2, Y dim 1
    (as in the reference's ba_interp2(ImgDef, v22, u22)); F is the MATLAB-shaped image."""
    if method != "cubic":
        raise ValueError("only 'cubic' is provided")
    ctx = get_context()
    F = np.asarray(F, dtype=np.float64)
    _ensure_images(F, F)
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    # the library samples at (x = dim-1 coordinate, y = dim-2 coordinate)
    Z = ctx.bicubic(Y.ravel(), X.ravel())
    return Z.reshape(X.shape)


# ---- integer search ---------------------------------------------------------------------

def funIntegerSearch(f, g, tempSizeOfSearchRegion, gridx, gridy, winsize, winstepsize,
                     tempNoOfInitPt=0, *varargin):
    """funIntegerSearch.m:3-80, whole-field mode (tempNoOfInitPt = 0)."""
    if tempNoOfInitPt != 0:
        raise NotImplementedError("seed-point mode (tempNoOfInitPt = 1) is interactive")
    s = np.atleast_1d(tempSizeOfSearchRegion).astype(int)
    if s.size == 1:
        s = np.array([s[0], s[0]])
    ctx = _ensure_images(f, g)
    grid = [int(gridx[0]), int(gridx[-1]), int(gridy[0]), int(gridy[-1])]
    r = ctx.integer_search(grid, int(winsize), int(winstepsize), s, qfactors=True)
    cc = dict(max=r["ccmax"], A=[], qfactors=r["qf"], status=r["status"])
    return r["x"], r["y"], r["u"], r["v"], cc


def IntegerSearch(ImgRef, ImgDef, file_name, DICpara):
    """IntegerSearch.m:1-167 without the prompts: the search radius comes from
    DICpara['SizeOfFFTSearchRegion']; masked nodes become NaN (:68-76); the median test
    of funRemoveOutliers(u, v, cc, 0.5, 100) runs (:113-114) but not the interactive
    cleanup that follows it."""
    s = DICpara["SizeOfFFTSearchRegion"]
    roi = DICpara["gridxyROIRange"]
    x0, y0, u, v, cc = funIntegerSearch(ImgRef, ImgDef, s, roi["gridx"], roi["gridy"],
                                        DICpara["winsize"], DICpara["winstepsize"], 0)
    mask = DICpara.get("ImgRefMask")
    if mask is not None:
        xi = x0.astype(int) - 1
        yi = y0.astype(int) - 1
        hole = np.asarray(mask)[xi, yi] == 0
        u = np.where(hole, np.nan, u)
        v = np.where(hole, np.nan, v)
    u, v, cc, _, _, _ = funRemoveOutliers(u, v, cc, 0.5, 100)
    DICpara = dict(DICpara)
    DICpara["InitFFTSearchMethod"] = 1
    return DICpara, x0, y0, u, v, cc


# ---- mesh and initial value --------------------------------------------------------------

def MeshSetUp(x, y, DICpara):
    """MeshSetUp.m:19-89: node = subset centre (x fastest), Q4 elements, empty BC lists.
    elementsFEM is returned 1-based like the reference; coordinatesFEM is (S, 2)."""
    x = np.asarray(x)
    y = np.asarray(y)
    N, M = x.shape  # N rows (y), M columns (x)
    coordinatesFEM = np.stack([x.ravel(), y.ravel()], axis=1)
    el = []
    for j in range(1, N):
        for i in range(1, M):
            el.append([(j - 1) * M + i, (j - 1) * M + i + 1, j * M + i + 1, j * M + i])
    elementsFEM = np.array(el, dtype=np.int64).reshape(-1, 4)
    ImgSize = DICpara.get("ImgSize", (0, 0))
    return dict(coordinatesFEM=coordinatesFEM, elementsFEM=elementsFEM, dirichlet=[],
                neumann=[], x0=x.T, y0=y.T, M=M, N=N,
                y0World=ImgSize[1] + 1 - y.T,
                coordinatesFEMWorld=np.stack([coordinatesFEM[:, 0],
                                              ImgSize[1] + 1 - coordinatesFEM[:, 1]], axis=1))


def inpaint_nans_springs(A):
    """inpaint_nans(A, 4) on the device (aldic_inpaint_nans): springs between each NaN and
    its horizontal and vertical neighbours, least squares with the known values fixed."""
    return get_context().inpaint_nans(A, 4)


def inpaint_nans_plate(A):
    """inpaint_nans(A, 3) on the device (aldic_inpaint_nans): the del^4 plate template,
    reduced near the border, least squares with the known values fixed."""
    return get_context().inpaint_nans(A, 3)


def funRemoveOutliers(u, v, cc, qDICOrNot, Thr0, *varargin):
    """funRemoveOutliers.m:1-86 without its prompts, on the device
    (aldic_remove_outliers): the single-precision normalised median test of u and v at
    threshold Thr0, the flagged elements and the NaNs in-painted by springs, the NaNs put
    back.  Returns (u, v, cc, BadptRow, BadptCol, RemoveOutliersList) with the list as
    MATLAB's 1-based column-major indices.  The q-factor removal (qDICOrNot == 1) and the
    interactive bound / click cleanup that follow when qDICOrNot != 0 are not provided;
    Thr0 must be given."""
    if qDICOrNot == 1:
        raise NotImplementedError("q-factor based removal (bimodal fit) is not provided")
    if not Thr0:
        raise NotImplementedError("the threshold prompt is interactive; pass Thr0")
    u2, v2, flagged = get_context().remove_outliers(u, v, Thr0)
    lst = np.flatnonzero(flagged.ravel(order="F")) + 1
    return u2, v2, cc, [], [], lst


def Init(u, v, Phi, x, y, index=0):
    """Init.m:1-179 with index = 0: inpaint NaNs (method 4) and interleave U0."""
    if index != 0:
        raise NotImplementedError("Init index 1 (outlier spring test) is not on the path")
    u = inpaint_nans_springs(u)
    v = inpaint_nans_springs(v)
    U0 = np.empty(2 * u.size)
    U0[0::2] = u.ravel()
    U0[1::2] = v.ravel()
    return U0


# ---- local steps -------------------------------------------------------------------------

def _refill(coordinatesFEM, U, F=None):
    """LocalICGN.m:141-161: scatteredInterpolant(good nodes, 'natural') of U (and F) at
    the NaN nodes, evaluated on the device (aldic_natural_fill)."""
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    bad = np.isnan(U[0::2]) | np.isnan(U[1::2])
    if F is not None:
        bad = bad | np.isnan(F.reshape(-1, 4)).any(axis=1)
    if not bad.any():
        return U, F
    return get_context().natural_fill(coords, bad, U, F)


def LocalICGN(U0, coordinatesFEM, Df, ImgRef, ImgDef, DICpara, ICGNmethod="GaussNewton",
              tol=1e-2):
    """LocalICGN.m:1-161 -> [U, F, HtempPar, LocalTime, ConvItPerEle, LocalICGNBadPtNum]."""
    if ICGNmethod != "GaussNewton":
        raise NotImplementedError("the reference's LM branch is dead code (funICGN.m:199)")
    ctx = _ensure_images(ImgRef, ImgDef, Df)
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    t0 = time.perf_counter()
    U, F, H, it, _ = ctx.icgn6(coords, U0, DICpara["winsize"], tol)
    LocalTime = time.perf_counter() - t0
    # bad points (:119-131)
    it = it.astype(np.int64)
    bad = (it < 0) | (it > 99)
    row3 = int(np.sum(it == 102))
    LocalICGNBadPtNum = int(np.sum(bad)) - row3
    good = ~bad
    if good.any():
        mean = it[good].mean()
        std = it[good].std(ddof=1) if good.sum() > 1 else 0.0
        bad = bad | (it > mean + 0.15 * std)
    U = U.copy()
    F = F.copy()
    U[0::2][bad] = np.nan
    U[1::2][bad] = np.nan
    for c in range(4):
        F[c::4][bad] = np.nan
    U, F = _refill(coords, U, F)
    return U, F, H, LocalTime, it, LocalICGNBadPtNum


def funICGN(U0, x0, y0, Df, ImgRef, ImgDef, winsize, tol, ICGNmethod="GaussNewton"):
    """funICGN.m:1-322 for one subset -> [U, F, stepwithinwhile, HGlobal]."""
    ctx = _ensure_images(ImgRef, ImgDef, Df)
    U, F, H, it, _ = ctx.icgn6(np.array([[x0, y0]]), np.asarray(U0, float)[:2], winsize, tol)
    return U, F, int(it[0]), H[0]


def Subpb1(UOld, FOld, udual, vdual, coordinatesFEM, Df, ImgRef, ImgDef, mu, beta, HPar,
           ALSolveStep, DICpara, ICGNmethod="GaussNewton", tol=1e-2, smoothing="gridfit"):
    """Subpb1.m:1-199 -> [U, HPar, ALSub1Time, ConvItPerEle, LocalICGNBadPtNum].

    HPar may be the reference's cell(21) (list of 21 node vectors) or an (S, 21) array.
    With ALSolveStep == 1 the reference passes a zero Hessian (:69-77).  Failed nodes
    (status < 0 or > 99) become NaN and u, v are then refitted at every node by gridfit
    with the springs regulariser over the node grid (:165-199, device CG); with
    smoothing='off' the failed nodes are refilled by the natural-neighbour interpolant
    instead (as LocalICGN does)."""
    if smoothing not in ("gridfit", "off"):
        raise ValueError("smoothing must be 'gridfit' or 'off'")
    ctx = _ensure_images(ImgRef, ImgDef, Df)
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    S = coords.shape[0]
    H = np.asarray(HPar, dtype=np.float64)
    if H.shape == (21, S):
        H = H.T
    if ALSolveStep <= 1:
        H = np.zeros((S, 21))
    t0 = time.perf_counter()
    U, it = ctx.icgn2(coords, DICpara["winsize"], H, mu, vdual, UOld, FOld, tol)
    ALSub1Time = time.perf_counter() - t0
    bad = (it < 0) | (it > 99)
    U = U.copy()
    U[0::2][bad] = np.nan
    U[1::2][bad] = np.nan
    if smoothing == "off":
        U, _ = _refill(coords, U)
    else:
        xs = np.unique(coords[:, 0])
        ys = np.unique(coords[:, 1])
        ix = np.searchsorted(xs, coords[:, 0])
        iy = np.searchsorted(ys, coords[:, 1])
        for c in range(2):
            z = np.full((ys.size, xs.size), np.nan)
            z[iy, ix] = U[c::2]
            zg, _ = ctx.gridfit(xs, ys, z)
            U[c::2] = zg[iy, ix]
    return U, HPar, ALSub1Time, it.astype(np.int64), int(bad.sum())


def funICGN_Subpb1(x0, y0, Df, ImgRef, ImgDef, winsize, H, beta, mu, udual, vdual, UOld, FOld,
                   tol, ICGNmethod="GaussNewton"):
    """funICGN_Subpb1.m:1-362 for one subset -> [U, H, stepwithinwhile]."""
    ctx = _ensure_images(ImgRef, ImgDef, Df)
    H = np.asarray(H, dtype=np.float64)
    h21 = H[np.tril_indices(6)[::-1]] if H.shape == (6, 6) else H.ravel()
    if H.shape == (6, 6):
        h21 = np.array([H[r, c] for c in range(6) for r in range(c, 6)])
    U, it = ctx.icgn2(np.array([[x0, y0]]), winsize, h21[None, :], mu,
                      np.asarray(vdual, float)[:2], np.asarray(UOld, float)[:2],
                      np.asarray(FOld, float)[:4], tol)
    return U, H, int(it[0])


# ---- global step -------------------------------------------------------------------------

_MESH_KEY = None


def _ensure_mesh(coordinatesFEM, elementsFEM):
    global _MESH_KEY
    ctx = get_context()
    key = _fingerprint(coordinatesFEM, elementsFEM)
    if key != _MESH_KEY:
        el = np.asarray(elementsFEM, dtype=np.int64)[:, :4] - 1
        ctx.set_mesh(np.asarray(coordinatesFEM, dtype=np.float64), el.astype(np.int32))
        _MESH_KEY = key
    return ctx


def Subpb2(DICmesh, beta, mu, U, F, udual, vdual, alpha=0.0, GaussPtOrder=2, rtol=1e-12):
    """Subpb2.m:1-197 -> Uhat (the sparse direct solve is replaced by FP64 Jacobi-PCG to
    relative residual rtol)."""
    ctx = _ensure_mesh(DICmesh["coordinatesFEM"], DICmesh["elementsFEM"])
    Uh, st = ctx.subpb2(beta, mu, U, F, udual, vdual, alpha=alpha, rtol=rtol)
    return Uh


def funGlobal_NodalStrainAvg(coordinatesFEM, elementsFEM, U, GaussPtOrder=2, swap_fix=False):
    """funGlobal_NodalStrainAvg.m:1-147 (reference slot order unless swap_fix)."""
    ctx = _ensure_mesh(coordinatesFEM, elementsFEM)
    return ctx.nodal_strain(U, swap_fix)


__all__ = ["AldicError", "get_context", "funNormalizeImg", "funImgGradient", "ba_interp2",
           "funIntegerSearch", "IntegerSearch", "MeshSetUp", "Init", "inpaint_nans_springs",
           "inpaint_nans_plate", "funRemoveOutliers",
           "LocalICGN", "funICGN", "Subpb1", "funICGN_Subpb1", "Subpb2",
           "funGlobal_NodalStrainAvg"]
