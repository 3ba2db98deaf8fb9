"""Image-sequence driver and the reference's result containers, on the device frame.

  main_ALDIC.m:40-43     ResultDisp, ResultDefGrad, ResultFEMeshEachFrame, ResultFEMesh
  main_ALDIC.m:55-163    per deformed frame: the integer search for the first (and with
                         NewFFTSearch for every) frame, else the previous frame's U as
                         the initial guess; ResultFEMesh{1} and ResultFEMeshEachFrame
  main_ALDIC.m:455-457   ResultDisp{k}.U, ResultDisp{k}.ALSub1BadPtNum, ResultDefGrad{k}.F
  main_ALDIC.m:536-620   Section 8 per frame: ComputeStrain.m on the frame's node grid

Cumulative mode only (every frame against the first image, the reference's
ImgSeqIncUnit > number of frames): the incremental reference update
(ReadImageRefUpdate, mesh moved with the solved displacements) is not built.  From the
sixth deformed frame on the reference predicts the initial guess with POD + Gaussian
process regression (funPOR_GPR.m, MATLAB's fitrgp with fitted hyperparameters); this
driver keeps the previous frame's U there, which changes only IC-GN's starting point.
"""
import numpy as np

from .aldic import get_context, MeshSetUp, AldicError


def _counts_local(iters6):
    """LocalICGN.m:117-122: statuses < 0 or > 99, less the masked ones (102)."""
    it = np.asarray(iters6)
    return int(np.sum((it < 0) | (it > 99)) - np.sum(it == 102))


def run_sequence(images, DICpara, mask=None, roi=None, ctx=None):
    """Runs every deformed frame of images = [ref, def_1, def_2, ...] (equal shape and
    dtype, raw pixels) against images[0] and returns the reference's containers:
    ResultDisp[k] = {U (2S), ALSub1BadPtNum (6)}, ResultDefGrad[k] = {F (4S)},
    ResultFEMesh[0] = {coordinatesFEM, elementsFEM (1-based), winsize, winstepsize,
    gridxyROIRange}, ResultFEMeshEachFrame[k] = {coordinatesFEM, elementsFEM}, plus DICmesh.

    DICpara keys: winsize, winstepsize, SizeOfFFTSearchRegion (int or [sx, sy]),
    gridxyROIRange ({gridx: (x1, x2), gridy: (y1, y2)}, 1-based; default whole image),
    tol (1e-2), mu (1e-3), beta (None: 1e-2 step^2 mu), ALSolveStep (ADMM steps, 3),
    NewFFTSearch (0), ImgSeqIncUnit (cumulative when absent or > number of frames)."""
    if len(images) < 2:
        raise ValueError("run_sequence: need the reference and at least one deformed image")
    nfr = len(images) - 1
    inc = DICpara.get("ImgSeqIncUnit", nfr + 1)
    if inc <= nfr:
        raise AldicError("run_sequence: incremental reference updates (ImgSeqIncUnit <= frames) "
                         "are not supported; cumulative mode only")
    ctx = ctx or get_context()
    f = np.asarray(images[0])
    ny, nx = f.shape
    ws, step = int(DICpara["winsize"]), int(DICpara["winstepsize"])
    s = DICpara.get("SizeOfFFTSearchRegion", 10)
    s = [int(s), int(s)] if np.isscalar(s) else [int(s[0]), int(s[1])]
    gr = DICpara.get("gridxyROIRange")
    grid = [1, nx, 1, ny] if gr is None else [int(gr["gridx"][0]), int(gr["gridx"][1]),
                                               int(gr["gridy"][0]), int(gr["gridy"][1])]
    n_admm = int(DICpara.get("ALSolveStep", 3))
    new_search = int(DICpara.get("NewFFTSearch", 0))
    ResultDisp, ResultDefGrad = [None] * nfr, [None] * nfr
    ResultFEMeshEachFrame = [None] * nfr
    ResultFEMesh = [None]
    DICmesh = None
    for k in range(nfr):
        g = np.asarray(images[k + 1])
        if g.shape != f.shape or g.dtype != f.dtype:
            raise ValueError("run_sequence: images differ in shape or dtype")
        ctx.frame_stage(f, g, mask=mask, roi=roi)
        S = ctx.frame_setup(ws, step, s, grid, tol=DICpara.get("tol", 1e-2),
                            mu=DICpara.get("mu", 1e-3), beta=DICpara.get("beta"),
                            n_admm=n_admm)
        if DICmesh is None:
            x1, y1, Nx, Ny = ctx.search_grid(grid, ws, step, s)
            xs = x1 + ws / 2.0 + step * np.arange(Nx)
            ys = y1 + ws / 2.0 + step * np.arange(Ny)
            X, Y = np.meshgrid(xs, ys)
            DICmesh = MeshSetUp(X, Y, dict(DICpara, ImgSize=(nx, ny)))
            assert DICmesh["coordinatesFEM"].shape[0] == S
            ResultFEMesh[0] = dict(coordinatesFEM=DICmesh["coordinatesFEM"],
                                   elementsFEM=DICmesh["elementsFEM"], winsize=ws,
                                   winstepsize=step,
                                   gridxyROIRange=dict(gridx=grid[:2], gridy=grid[2:]))
        if k == 0 or new_search:
            ctx.frame_run()
        else:
            ctx.frame_run_seeded(ResultDisp[k - 1]["U"])
        out = ctx.frame_get()
        cnt = ctx.frame_counters()["subpb1_bad"]
        bad = np.zeros(6)
        bad[0] = _counts_local(out["iters6"])
        for st in range(min(n_admm, 5)):
            bad[st + 1] = cnt[st] if cnt[st] >= 0 else 0
        ResultDisp[k] = dict(U=out["U"].copy(), ALSub1BadPtNum=bad)
        ResultDefGrad[k] = dict(F=out["F"].copy())
        ResultFEMeshEachFrame[k] = dict(coordinatesFEM=DICmesh["coordinatesFEM"],
                                        elementsFEM=DICmesh["elementsFEM"])
    return dict(ResultDisp=ResultDisp, ResultDefGrad=ResultDefGrad, ResultFEMesh=ResultFEMesh,
                ResultFEMeshEachFrame=ResultFEMeshEachFrame, DICmesh=DICmesh)


def ComputeStrain(DICpara, ULocal, FLocal, M, N, USubpb2=None, coordinatesFEM=None,
                  elementsFEM=None, Rad=None, ctx=None):
    """ComputeStrain.m:11-139 on the device for one frame's M x N node grid (x fastest).
    DICpara: MethodToComputeStrain (0 solved F, 1 finite differences, 2 plane fit of
    half width Rad, 3 finite-element nodal strains of USubpb2 on the mesh), StrainType
    (0 infinitesimal, 1 Eulerian, 2 Green-Lagrangian), winstepsize.  Returns
    FStrainFinite (4MN), FStraintemp and FStrainWorld (the nodes at least Rad from the
    border; F21, F12 negated in the world copy), Rad and temp3 (0-based entries)."""
    own = ctx is None
    ctx = ctx or get_context()
    method = int(DICpara["MethodToComputeStrain"])
    stype = int(DICpara.get("StrainType", 0))
    h = float(DICpara["winstepsize"])
    if method == 2 and Rad is None:
        raise ValueError("ComputeStrain: the plane fit needs its half window Rad")
    if method == 3:
        if own:
            from .aldic import _ensure_mesh  # the default context caches its mesh
            _ensure_mesh(coordinatesFEM, elementsFEM)
        else:
            el = np.asarray(elementsFEM, dtype=np.int64)[:, :4] - 1  # 1-based elements
            ctx.set_mesh(np.asarray(coordinatesFEM, dtype=np.float64), el.astype(np.int32))
        U = USubpb2
    else:
        U = ULocal
    Ff, Fw, r = ctx.compute_strain(M, N, h, U=U, F=FLocal, method=method,
                                   rad=int(Rad) if Rad is not None else 1, strain_type=stype)
    t2 = np.arange(M * N).reshape(N, M).T[r:M - r, r:N - r].flatten(order="F")
    temp3 = (4 * t2[:, None] + np.arange(4)[None, :]).ravel()
    Ft = Fw.copy()
    Ft[1::4] = -Ft[1::4]
    Ft[2::4] = -Ft[2::4]
    return dict(FStrainFinite=Ff, FStraintemp=Ft, FStrainWorld=Fw, Rad=r, temp3=temp3)


__all__ = ["run_sequence", "ComputeStrain"]
