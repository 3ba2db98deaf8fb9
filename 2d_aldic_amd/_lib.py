"""ctypes binding of lib/libaldic.so (the C ABI declared in include/aldic.h).

The library must have been built (``python 2d_aldic_amd/build.py``); loading fails
loudly otherwise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libaldic.so")

ALDIC_U8, ALDIC_U16, ALDIC_F32, ALDIC_F64 = 0, 1, 2, 3
_DTYPES = {np.dtype(np.uint8): ALDIC_U8, np.dtype(np.uint16): ALDIC_U16,
           np.dtype(np.float32): ALDIC_F32, np.dtype(np.float64): ALDIC_F64}

EXPORTS = (
    "aldic_create", "aldic_destroy", "aldic_last_error", "aldic_set_images",
    "aldic_set_normalized", "aldic_set_gradients", "aldic_get_images", "aldic_search_grid",
    "aldic_integer_search", "aldic_icgn6", "aldic_icgn2", "aldic_set_mesh", "aldic_subpb2",
    "aldic_nodal_strain", "aldic_bicubic", "aldic_frame_stage", "aldic_frame_setup",
    "aldic_frame_run", "aldic_frame_get", "aldic_frame_times", "aldic_gridfit",
    "aldic_frame_counters", "aldic_frame_scalars", "aldic_set_quadtree",
    "aldic_subpb2_quadtree", "aldic_nodal_strain_quadtree", "aldic_qframe_setup",
    "aldic_qframe_run", "aldic_qframe_get", "aldic_prep_noalloc", "aldic_natural_fill",
    "aldic_frame_run_seeded", "aldic_compute_strain", "aldic_inpaint_nans",
    "aldic_remove_outliers", "aldic_icgn6_start", "aldic_cnn_nparams", "aldic_cnn_set_weights",
    "aldic_cnn_get_weights", "aldic_cnn_predict", "aldic_cnn_add_samples",
    "aldic_cnn_clear_samples", "aldic_cnn_train", "aldic_cnn_grad", "aldic_local_icgn",
    "aldic_integer_search_mg", "aldic_smooth_strain_quadtree",
)


class FrameOpts(C.Structure):
    _fields_ = [("ws", C.c_int), ("step", C.c_int), ("s", C.c_int * 2), ("grid", C.c_int * 4),
                ("tol", C.c_double), ("mu", C.c_double), ("beta", C.c_double),
                ("n_admm", C.c_int), ("cg_maxit", C.c_int), ("cg_rtol", C.c_double),
                ("no_subpb1_fit", C.c_int), ("stage_times", C.c_int), ("smooth", C.c_int),
                ("beta_sweep", C.c_int), ("admm_tol", C.c_double), ("outlier_thr", C.c_double),
                ("learned_start", C.c_int)]


class QFrameOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("mu", C.c_double), ("beta", C.c_double),
                ("n_admm", C.c_int), ("cg_maxit", C.c_int), ("cg_rtol", C.c_double),
                ("smooth_h", C.c_double), ("smoothness", C.c_double)]


_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run 2d_aldic_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    P = C.c_void_p
    L.aldic_create.argtypes = [C.POINTER(P), ip]
    L.aldic_destroy.argtypes = [P]
    L.aldic_destroy.restype = None
    L.aldic_last_error.argtypes = [P]
    L.aldic_last_error.restype = C.c_char_p
    L.aldic_set_images.argtypes = [P, vp, vp, ip, ip, ip, vp, vp]
    L.aldic_set_normalized.argtypes = [P, vp, vp, ip, ip]
    L.aldic_set_gradients.argtypes = [P, vp, vp]
    L.aldic_get_images.argtypes = [P, vp, vp, vp, vp]
    L.aldic_search_grid.argtypes = [ip, ip, vp, ip, ip, vp, vp]
    L.aldic_integer_search.argtypes = [P, vp, ip, ip, vp, vp, vp, vp, vp, vp, vp, vp]
    L.aldic_smooth_strain_quadtree.argtypes = [P, ip, vp, dp, dp, dp, ip, vp, vp]
    L.aldic_integer_search_mg.argtypes = [P, vp, ip, ip, C.c_long, vp, vp, vp, vp, vp, vp,
                                          C.c_long, vp]
    L.aldic_icgn6.argtypes = [P, ip, vp, vp, vp, ip, dp, vp, vp, vp, vp, vp]
    L.aldic_icgn2.argtypes = [P, ip, vp, vp, ip, vp, dp, vp, vp, vp, dp, vp, vp]
    L.aldic_set_mesh.argtypes = [P, vp, ip, vp, ip]
    L.aldic_subpb2.argtypes = [P, dp, dp, dp, vp, vp, vp, vp, dp, ip, vp, vp]
    L.aldic_nodal_strain.argtypes = [P, vp, ip, vp]
    L.aldic_bicubic.argtypes = [P, vp, vp, C.c_long, vp]
    L.aldic_frame_stage.argtypes = [P, vp, vp, ip, ip, ip, vp, vp]
    L.aldic_frame_setup.argtypes = [P, C.POINTER(FrameOpts), C.POINTER(ip)]
    L.aldic_frame_run.argtypes = [P]
    L.aldic_frame_get.argtypes = [P, vp, vp, vp, vp, vp, vp, vp]
    L.aldic_frame_times.argtypes = [P, vp, ip]
    L.aldic_gridfit.argtypes = [P, ip, ip, vp, vp, vp, dp, ip, vp, vp]
    L.aldic_frame_counters.argtypes = [P, vp, ip]
    L.aldic_frame_scalars.argtypes = [P, vp, ip]
    L.aldic_set_quadtree.argtypes = [P, vp, ip, vp, ip, vp, ip]
    L.aldic_subpb2_quadtree.argtypes = [P, dp, dp, vp, vp, dp, ip, vp, vp]
    L.aldic_nodal_strain_quadtree.argtypes = [P, vp, vp, vp, vp, vp, vp]
    L.aldic_qframe_setup.argtypes = [P, C.POINTER(QFrameOpts), vp, vp, vp]
    L.aldic_qframe_run.argtypes = [P, vp]
    L.aldic_qframe_get.argtypes = [P, vp, vp, vp, vp, vp]
    L.aldic_prep_noalloc.argtypes = [P, vp, vp, ip, ip, ip, vp, vp, vp, vp, vp, vp, vp]
    L.aldic_natural_fill.argtypes = [P, ip, vp, vp, vp, vp]
    L.aldic_frame_run_seeded.argtypes = [P, vp]
    L.aldic_compute_strain.argtypes = [P, ip, ip, dp, vp, vp, ip, ip, ip, vp, vp]
    L.aldic_inpaint_nans.argtypes = [P, vp, ip, ip, ip]
    L.aldic_remove_outliers.argtypes = [P, vp, vp, ip, ip, dp, vp, vp]
    L.aldic_icgn6_start.argtypes = [P, ip, vp, vp, vp, ip, dp, vp, vp, vp, vp, vp]
    L.aldic_cnn_nparams.argtypes = []
    L.aldic_cnn_nparams.restype = C.c_long
    L.aldic_cnn_set_weights.argtypes = [P, vp, C.c_long]
    L.aldic_cnn_get_weights.argtypes = [P, vp, C.c_long]
    L.aldic_cnn_predict.argtypes = [P, ip, vp, vp]
    L.aldic_cnn_add_samples.argtypes = [P, ip, vp, vp]
    L.aldic_cnn_clear_samples.argtypes = [P]
    L.aldic_cnn_train.argtypes = [P, ip, ip, C.c_float, C.c_ulonglong, vp]
    L.aldic_cnn_grad.argtypes = [P, ip, vp, vp, vp, vp]
    L.aldic_local_icgn.argtypes = [P, ip, vp, ip, ip, vp, C.c_double, ip, C.c_double,
                                   vp, vp, vp, vp, vp, vp, vp, vp, vp]
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class AldicError(RuntimeError):
    pass


class Context:
    """One library context on one HIP device; owns the device-side state."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.aldic_create(C.byref(h), int(device))
        if rc != 0:
            raise AldicError(f"aldic_create failed ({rc}); is a GPU visible?")
        self.h = h
        self.nx = self.ny = 0
        self.n_nodes = 0
        self.n_quad_nodes = 0
        self.n_quad_elems = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.aldic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            msg = self.L.aldic_last_error(self.h).decode(errors="replace")
            raise AldicError(f"{what} failed ({rc}): {msg}")

    # -- images --------------------------------------------------------------------
    def set_images(self, f, g, mask=None, roi=None):
        """Raw pair, (ny, nx) C-ordered (x fastest), uint8/uint16/float32/float64."""
        f = np.ascontiguousarray(f)
        g = np.ascontiguousarray(g, dtype=f.dtype)
        ny, nx = f.shape
        m = None if mask is None else _c(mask, np.uint8)
        r = None if roi is None else _c(roi, np.int32)
        self._ck(self.L.aldic_set_images(self.h, _p(f), _p(g), nx, ny, _DTYPES[f.dtype], _p(m),
                                         _p(r)), "aldic_set_images")
        self.nx, self.ny = nx, ny

    def set_normalized(self, f, g):
        f = _c(f, np.float64)
        g = _c(g, np.float64)
        ny, nx = f.shape
        self._ck(self.L.aldic_set_normalized(self.h, _p(f), _p(g), nx, ny), "aldic_set_normalized")
        self.nx, self.ny = nx, ny

    def set_gradients(self, gx, gy):
        gx = _c(gx, np.float64)
        gy = _c(gy, np.float64)
        self._ck(self.L.aldic_set_gradients(self.h, _p(gx), _p(gy)), "aldic_set_gradients")

    def get_images(self):
        out = [np.empty((self.ny, self.nx)) for _ in range(4)]
        self._ck(self.L.aldic_get_images(self.h, *[_p(o) for o in out]), "aldic_get_images")
        return out

    # -- search ---------------------------------------------------------------------
    def search_grid(self, grid, ws, step, s):
        out = np.zeros(4, np.int32)
        rc = self.L.aldic_search_grid(self.nx, self.ny, _p(_c(grid, np.int32)), int(ws),
                                      int(step), _p(_c(s, np.int32)), _p(out))
        if rc != 0:
            raise AldicError("aldic_search_grid: bad arguments")
        return [int(v) for v in out]

    def integer_search(self, grid, ws, step, s, qfactors=False):
        grid = _c(grid, np.int32)
        s = _c(s, np.int32)
        x1, y1, Nx, Ny = self.search_grid(grid, ws, step, s)
        S = Nx * Ny
        outs = [np.empty(S) for _ in range(5)]
        qf = np.empty((S, 2)) if qfactors else None
        st = np.empty(S, np.int32)
        self._ck(self.L.aldic_integer_search(self.h, _p(grid), int(ws), int(step), _p(s),
                                             *[_p(o) for o in outs], _p(qf), _p(st)),
                 "aldic_integer_search")
        x, y, u, v, cm = [o.reshape(Ny, Nx) for o in outs]
        return dict(x=x, y=y, u=u, v=v, ccmax=cm, qf=qf, status=st.reshape(Ny, Nx), x1=x1,
                    y1=y1, Nx=Nx, Ny=Ny)

    def smooth_strain_quadtree(self, coords, F, h, smoothness=1e-4, filter_std=0.5, filter_size=0):
        """funSmoothStrainQuadtree(F, DICmesh, DICpara) on the device: (smoothed F, CG
        iterations of the two passes)."""
        coords = _c(coords, np.float64).reshape(-1, 2)
        n = coords.shape[0]
        F = _c(F, np.float64).copy()
        assert F.size == 4 * n
        it = np.zeros(2, np.int32)
        self._ck(self.L.aldic_smooth_strain_quadtree(self.h, n, _p(coords), float(h), float(smoothness),
                                                     float(filter_std), int(filter_size), _p(F), _p(it)),
                 "aldic_smooth_strain_quadtree")
        return F, it

    def integer_search_mg(self, grid, ws, step):
        """funIntegerSearchMg over grid = (x1, x2, y1, y2): dict(x, y, u, v, ccmax) on the
        fitted (Ny, Nx) grid, the final regions' tempx, tempy, tempu, tempv, tempphi, and
        levels, n_nan (regions the reference's catch would give NaN), fit_iters."""
        grid = _c(grid, np.int32)
        cap = ((int(grid[1]) - int(grid[0])) // int(step) + 1) * \
              ((int(grid[3]) - int(grid[2])) // int(step) + 1)
        cap = max(cap, 4)
        ncap = 4 * (2 * (int(grid[1]) - int(grid[0])) // int(ws) + 2) * \
                   (2 * (int(grid[3]) - int(grid[2])) // int(ws) + 2)
        dims = np.zeros(6, np.int32)
        for _ in range(2):
            outs = [np.empty(cap) for _ in range(5)]
            temp = np.empty(5 * ncap)
            rc = self.L.aldic_integer_search_mg(self.h, _p(grid), int(ws), int(step), cap,
                                                _p(dims), *[_p(o) for o in outs], temp.size,
                                                _p(temp))
            if rc != 0 and 5 * int(dims[3]) > temp.size and int(dims[3]) > 0:
                ncap = int(dims[3])
                continue
            self._ck(rc, "aldic_integer_search_mg")
            break
        Nx, Ny, nr = int(dims[0]), int(dims[1]), int(dims[3])
        x, y, u, v, cm = [o[:Nx * Ny].reshape(Ny, Nx) for o in outs]
        t = temp[:5 * nr].reshape(5, nr)
        return dict(x=x, y=y, u=u, v=v, ccmax=cm, tempx=t[0], tempy=t[1], tempu=t[2],
                    tempv=t[3], tempphi=t[4], levels=int(dims[2]), n_nan=int(dims[4]),
                    fit_iters=int(dims[5]), Nx=Nx, Ny=Ny)

    # -- IC-GN -----------------------------------------------------------------------
    def icgn6(self, coords, U0, ws, tol=1e-2):
        coords = _c(coords, np.float64).reshape(-1, 2)
        S = coords.shape[0]
        U0 = _c(U0, np.float64)
        wsa = None if np.isscalar(ws) else _c(ws, np.int32)
        U = np.empty(2 * S)
        F = np.empty(4 * S)
        H = np.empty((S, 21))
        it = np.empty(S, np.int32)
        z = np.empty(S)
        self._ck(self.L.aldic_icgn6(self.h, S, _p(coords), _p(U0), _p(wsa),
                                    int(ws) if wsa is None else 0, float(tol), _p(U), _p(F),
                                    _p(H), _p(it), _p(z)), "aldic_icgn6")
        return U, F, H, it, z

    def icgn6_start(self, coords, P0, ws, tol=1e-2):
        """6-DOF IC-GN started from P0 (S x 6: u_x, v_x, u_y, v_y, u, v)."""
        coords = _c(coords, np.float64).reshape(-1, 2)
        S = coords.shape[0]
        P0 = _c(P0, np.float64).reshape(S, 6)
        wsa = None if np.isscalar(ws) else _c(ws, np.int32)
        U = np.empty(2 * S)
        F = np.empty(4 * S)
        H = np.empty((S, 21))
        it = np.empty(S, np.int32)
        z = np.empty(S)
        self._ck(self.L.aldic_icgn6_start(self.h, S, _p(coords), _p(P0), _p(wsa),
                                          int(ws) if wsa is None else 0, float(tol), _p(U),
                                          _p(F), _p(H), _p(it), _p(z)), "aldic_icgn6_start")
        return U, F, H, it, z

    # ---- learned initial guess ----------------------------------------------------
    def cnn_nparams(self):
        return int(self.L.aldic_cnn_nparams())

    def cnn_set_weights(self, w):
        w = _c(w, np.float32).ravel()
        self._ck(self.L.aldic_cnn_set_weights(self.h, _p(w), w.size), "aldic_cnn_set_weights")

    def cnn_get_weights(self):
        w = np.empty(self.cnn_nparams(), np.float32)
        self._ck(self.L.aldic_cnn_get_weights(self.h, _p(w), w.size), "aldic_cnn_get_weights")
        return w

    def cnn_predict(self, coords):
        coords = _c(coords, np.float64).reshape(-1, 2)
        P = np.empty((coords.shape[0], 6))
        self._ck(self.L.aldic_cnn_predict(self.h, coords.shape[0], _p(coords), _p(P)),
                 "aldic_cnn_predict")
        return P

    def cnn_add_samples(self, coords, P):
        coords = _c(coords, np.float64).reshape(-1, 2)
        P = _c(P, np.float64).reshape(coords.shape[0], 6)
        self._ck(self.L.aldic_cnn_add_samples(self.h, coords.shape[0], _p(coords), _p(P)),
                 "aldic_cnn_add_samples")

    def cnn_clear_samples(self):
        self._ck(self.L.aldic_cnn_clear_samples(self.h), "aldic_cnn_clear_samples")

    def cnn_train(self, steps, batch, lr, seed=0):
        loss = np.empty(max(int(steps), 1))
        self._ck(self.L.aldic_cnn_train(self.h, int(steps), int(batch), float(lr), int(seed),
                                        _p(loss)), "aldic_cnn_train")
        return loss[:int(steps)]

    def cnn_grad(self, rows):
        rows = _c(rows, np.int32)
        g = np.empty(self.cnn_nparams(), np.float32)
        loss = np.empty(1)
        out = np.empty((rows.size, 6))
        self._ck(self.L.aldic_cnn_grad(self.h, rows.size, _p(rows), _p(g), _p(loss), _p(out)),
                 "aldic_cnn_grad")
        return g, float(loss[0]), out

    def local_icgn(self, grid, ws, step, s, start="search", tol=1e-2, max_it=20, min_zncc=0.9):
        """LocalICGN's start and 6-DOF IC-GN over the whole-field search grid, on the device:
        start "search" (the integer search) or "learned" (the network's warp, the search as
        fallback).  Returns dict(x, y, U, F, iters, zncc, fallback, n_fallback, times_ms)."""
        grid = _c(grid, np.int32)
        sv = _c(np.broadcast_to(np.asarray(s), (2,)), np.int32)
        g4 = self.search_grid(grid, ws, step, sv)
        S = g4[2] * g4[3]
        x, y, zz = np.empty(S), np.empty(S), np.empty(S)
        U, F = np.empty(2 * S), np.empty(4 * S)
        it, fb = np.empty(S, np.int32), np.empty(S, np.int32)
        nfb = np.zeros(1, np.int32)
        ms = np.zeros(5, np.float32)
        st = {"search": 0, "learned": 1}[start]
        self._ck(self.L.aldic_local_icgn(self.h, st, _p(grid), int(ws), int(step), _p(sv), float(tol),
                                         int(max_it), float(min_zncc), _p(x), _p(y), _p(U), _p(F),
                                         _p(it), _p(zz), _p(fb), _p(nfb), _p(ms)), "aldic_local_icgn")
        return dict(x=x, y=y, U=U, F=F, iters=it, zncc=zz, fallback=fb.astype(bool),
                    n_fallback=int(nfb[0]), times_ms=ms.astype(float), Nx=g4[2], Ny=g4[3])

    def icgn2(self, coords, ws, H21, mu, vdual, Uold, Fold, tol=1e-2):
        coords = _c(coords, np.float64).reshape(-1, 2)
        S = coords.shape[0]
        wsa = None if np.isscalar(ws) else _c(ws, np.int32)
        U = np.empty(2 * S)
        it = np.empty(S, np.int32)
        self._ck(self.L.aldic_icgn2(self.h, S, _p(coords), _p(wsa), int(ws) if wsa is None else 0,
                                    _p(_c(H21, np.float64)), float(mu), _p(_c(vdual, np.float64)),
                                    _p(_c(Uold, np.float64)), _p(_c(Fold, np.float64)),
                                    float(tol), _p(U), _p(it)), "aldic_icgn2")
        return U, it

    # -- FEM ---------------------------------------------------------------------------
    def set_mesh(self, coords, elems):
        coords = _c(coords, np.float64).reshape(-1, 2)
        elems = _c(elems, np.int32).reshape(-1, 4)
        self._ck(self.L.aldic_set_mesh(self.h, _p(coords), coords.shape[0], _p(elems),
                                       elems.shape[0]), "aldic_set_mesh")
        self.n_nodes = coords.shape[0]

    def subpb2(self, beta, mu, U, F, udual, vdual, alpha=0.0, rtol=1e-12, maxit=5000):
        n = self.n_nodes
        Uh = np.empty(2 * n)
        st = np.zeros(2)
        self._ck(self.L.aldic_subpb2(self.h, float(beta), float(mu), float(alpha),
                                     _p(_c(U, np.float64)), _p(_c(F, np.float64)),
                                     _p(_c(udual, np.float64)), _p(_c(vdual, np.float64)),
                                     float(rtol), int(maxit), _p(Uh), _p(st)), "aldic_subpb2")
        return Uh, st

    def set_quadtree(self, coords, elems8, dirichlet=()):
        """Quadtree mesh: coords (nn, 2), elems8 (E, 8) 0-based with -1 for an absent
        midside node, Dirichlet node indices (0-based)."""
        coords = _c(coords, np.float64).reshape(-1, 2)
        elems8 = _c(elems8, np.int32).reshape(-1, 8)
        d = _c(dirichlet, np.int32).ravel()
        self._ck(self.L.aldic_set_quadtree(self.h, _p(coords), coords.shape[0], _p(elems8),
                                           elems8.shape[0], _p(d) if d.size else None, d.size),
                 "aldic_set_quadtree")
        self.n_quad_nodes = coords.shape[0]
        self.n_quad_elems = elems8.shape[0]

    def subpb2_quadtree(self, beta, mu, U, F, rtol=1e-12, maxit=5000):
        n = self.n_quad_nodes
        Uh = np.empty(2 * n)
        st = np.zeros(2)
        self._ck(self.L.aldic_subpb2_quadtree(self.h, float(beta), float(mu),
                                              _p(_c(U, np.float64)), _p(_c(F, np.float64)),
                                              float(rtol), int(maxit), _p(Uh), _p(st)),
                 "aldic_subpb2_quadtree")
        return Uh, st

    def nodal_strain_quadtree(self, U, tri=None, w=None):
        """Gauss-point strains of the quadtree mesh (4E x [u_x u_y v_x v_y]) and their
        coordinates (4E x 2); with interpolation weights (tri, w: nn x 3) also the nodal
        values (4 nn, interleaved per node), else None."""
        n, ngp = self.n_quad_nodes, 4 * self.n_quad_elems
        Fgp = np.empty((ngp, 4))
        xy = np.empty((ngp, 2))
        F = None
        if tri is not None:
            tri = _c(tri, np.int32).reshape(n, 3)
            w = _c(w, np.float64).reshape(n, 3)
            F = np.empty(4 * n)
        self._ck(self.L.aldic_nodal_strain_quadtree(self.h, _p(_c(U, np.float64)), _p(tri), _p(w),
                                                    _p(F), _p(Fgp), _p(xy)),
                 "aldic_nodal_strain_quadtree")
        return F, Fgp, xy

    def qframe_setup(self, ws, tri, w, tol=1e-2, mu=1e-3, beta=None, n_admm=3, step=None,
                     cg_maxit=5000, cg_rtol=1e-10, smooth_h=0.0, smoothness=1e-4):
        """Quadtree frame on the mesh of set_quadtree: per-node windows ws (LocalICGN and
        Subpb1), Gauss-point weights tri / w; beta defaults to 1e-3 step^2 mu
        (main_ALDIC_Quadtree.m:202) with step = mean(ws) / 2."""
        ws = _c(ws, np.int32).ravel()
        if beta is None:
            st = float(np.mean(ws)) / 2.0 if step is None else float(step)
            beta = 1e-3 * st ** 2 * mu
        o = QFrameOpts(float(tol), float(mu), float(beta), int(n_admm), int(cg_maxit),
                       float(cg_rtol), float(smooth_h), float(smoothness))
        self._ck(self.L.aldic_qframe_setup(self.h, C.byref(o), _p(ws),
                                           _p(_c(tri, np.int32)), _p(_c(w, np.float64))),
                 "aldic_qframe_setup")
        self.qframe_opts = o

    def qframe_run(self, U0):
        self._ck(self.L.aldic_qframe_run(self.h, _p(_c(U0, np.float64))), "aldic_qframe_run")

    def qframe_get(self):
        n = self.n_quad_nodes
        U = np.empty(2 * n)
        F = np.empty(4 * n)
        it6 = np.empty(n, np.int32)
        it2 = np.empty(n, np.int32)
        t = np.empty(6)
        self._ck(self.L.aldic_qframe_get(self.h, _p(U), _p(F), _p(it6), _p(it2), _p(t)),
                 "aldic_qframe_get")
        return dict(U=U, F=F, iters6=it6, iters2=it2, frame_ms=t[0], local_ms=t[1],
                    global_ms=t[2], pcg_iters=int(t[3]), pcg_budget=int(t[4]),
                    pcg_reruns=int(t[5]))

    def nodal_strain(self, U, swap_fix=False):
        F = np.empty(4 * self.n_nodes)
        self._ck(self.L.aldic_nodal_strain(self.h, _p(_c(U, np.float64)), int(bool(swap_fix)),
                                           _p(F)), "aldic_nodal_strain")
        return F

    def inpaint_nans(self, A, method):
        """inpaint_nans(A, method) for method 3 (plate) or 4 (springs) on a 2-D array."""
        A = np.asarray(A, dtype=np.float64)
        rows, cols = A.shape
        f = np.asfortranarray(A).copy(order="F")
        self._ck(self.L.aldic_inpaint_nans(self.h, f.ctypes.data, rows, cols, int(method)),
                 "aldic_inpaint_nans")
        return np.ascontiguousarray(f)

    def remove_outliers(self, u, v, thr):
        """funRemoveOutliers(u, v, [], 0.5, thr) without prompts -> (u, v, flagged)."""
        u = np.asarray(u, dtype=np.float64)
        rows, cols = u.shape
        uf = np.asfortranarray(u).copy(order="F")
        vf = np.asfortranarray(np.asarray(v, dtype=np.float64)).copy(order="F")
        fl = np.zeros((rows, cols), np.int32, order="F")
        nf = C.c_int(0)
        self._ck(self.L.aldic_remove_outliers(self.h, uf.ctypes.data, vf.ctypes.data, rows, cols,
                                              float(thr), fl.ctypes.data, C.byref(nf)),
                 "aldic_remove_outliers")
        return np.ascontiguousarray(uf), np.ascontiguousarray(vf), fl.astype(bool)

    def compute_strain(self, M, N, h, U=None, F=None, method=1, rad=1, strain_type=0):
        """ComputeStrain.m on an M x N node grid -> (FStrainFinite (4 M N), FStrainWorld
        (the trimmed nodes, F21 and F12 negated), r).  Method 3 uses the mesh of
        set_mesh."""
        r = 0 if method == 0 else (rad if method == 2 else 1)
        nw = max(M - 2 * r, 0) * max(N - 2 * r, 0)
        Ff = np.empty(4 * M * N)
        Fw = np.empty(4 * nw) if nw else None
        self._ck(self.L.aldic_compute_strain(
            self.h, int(M), int(N), float(h),
            _p(None if U is None else _c(U, np.float64)), _p(None if F is None else _c(F, np.float64)),
            int(method), int(rad), int(strain_type), _p(Ff), _p(Fw)), "aldic_compute_strain")
        return Ff, (Fw if Fw is not None else np.empty(0)), r

    def natural_fill(self, coords, bad, U, F=None):
        """scatteredInterpolant(good nodes, 'natural') of U (2 per node) and F (4 per node)
        at the bad nodes (LocalICGN.m:141-161) -> refilled copies (U, F)."""
        coords = _c(coords, np.float64).reshape(-1, 2)
        n = coords.shape[0]
        b = _c(np.asarray(bad, bool).astype(np.int32), np.int32)
        U = _c(U, np.float64).copy()
        F = None if F is None else _c(F, np.float64).copy()
        assert U.size == 2 * n and (F is None or F.size == 4 * n)
        self._ck(self.L.aldic_natural_fill(self.h, n, _p(coords), _p(b), _p(U), _p(F)),
                 "aldic_natural_fill")
        return U, F

    def bicubic(self, X, Y):
        X = _c(X, np.float64).ravel()
        Y = _c(Y, np.float64).ravel()
        Z = np.empty_like(X)
        self._ck(self.L.aldic_bicubic(self.h, _p(X), _p(Y), X.size, _p(Z)), "aldic_bicubic")
        return Z

    def gridfit(self, xnodes, ynodes, z, rtol=1e-12, maxit=4000):
        """gridfit.m with Subpb1's options for data on the nodes: z (ny, nx) with NaN
        where there is no datum -> zgrid (ny, nx), and {iterations, relative residual}."""
        xnodes = _c(xnodes, np.float64).ravel()
        ynodes = _c(ynodes, np.float64).ravel()
        z = _c(z, np.float64).reshape(ynodes.size, xnodes.size)
        out = np.empty_like(z)
        st = np.zeros(2)
        self._ck(self.L.aldic_gridfit(self.h, xnodes.size, ynodes.size, _p(xnodes), _p(ynodes),
                                      _p(z), C.c_double(rtol), int(maxit), _p(out), _p(st)),
                 "aldic_gridfit")
        return out, st

    # -- frame ---------------------------------------------------------------------------
    def frame_stage(self, f, g, mask=None, roi=None):
        f = np.ascontiguousarray(f)
        g = np.ascontiguousarray(g, dtype=f.dtype)
        ny, nx = f.shape
        m = None if mask is None else _c(mask, np.uint8)
        r = None if roi is None else _c(roi, np.int32)
        self._ck(self.L.aldic_frame_stage(self.h, _p(f), _p(g), nx, ny, _DTYPES[f.dtype], _p(m),
                                          _p(r)), "aldic_frame_stage")
        self.nx, self.ny = nx, ny

    def frame_setup(self, ws, step, s, grid, tol=1e-2, mu=1e-3, beta=None, n_admm=3,
                    cg_maxit=2000, cg_rtol=1e-10, subpb1_fit=True, stage_times=True,
                    smooth=False, beta_sweep=False, admm_tol=0.0, outlier_thr=100.0,
                    learned_start=False):
        o = FrameOpts()
        o.ws, o.step = int(ws), int(step)
        o.s[0], o.s[1] = (int(s), int(s)) if np.isscalar(s) else (int(s[0]), int(s[1]))
        for k in range(4):
            o.grid[k] = int(grid[k])
        o.tol = tol
        o.mu = mu
        o.beta = beta if beta is not None else 1e-2 * step * step * mu
        o.n_admm = n_admm
        o.cg_maxit = cg_maxit
        o.cg_rtol = cg_rtol
        o.no_subpb1_fit = 0 if subpb1_fit else 1
        o.stage_times = 1 if stage_times else 0
        o.smooth = 1 if smooth else 0
        o.beta_sweep = 1 if beta_sweep else 0
        o.admm_tol = float(admm_tol)
        o.outlier_thr = float(outlier_thr)
        o.learned_start = 1 if learned_start else 0
        n = C.c_int(0)
        self._ck(self.L.aldic_frame_setup(self.h, C.byref(o), C.byref(n)), "aldic_frame_setup")
        self.frame_nodes = n.value
        self.frame_opts = o
        return n.value

    def frame_run(self):
        self._ck(self.L.aldic_frame_run(self.h), "aldic_frame_run")

    def frame_run_seeded(self, U0):
        """The frame from the initial guess U0 (2S) instead of the integer search (a later
        frame of a sequence, main_ALDIC.m:134-137)."""
        U0 = _c(U0, np.float64)
        assert U0.size == 2 * self.frame_nodes
        self._ck(self.L.aldic_frame_run_seeded(self.h, _p(U0)), "aldic_frame_run_seeded")

    def frame_get(self):
        S = self.frame_nodes
        U, F, U6, F6 = np.empty(2 * S), np.empty(4 * S), np.empty(2 * S), np.empty(4 * S)
        i6, i2 = np.empty(S, np.int32), np.empty(S, np.int32)
        u0 = np.empty(2 * S)
        self._ck(self.L.aldic_frame_get(self.h, _p(U), _p(F), _p(U6), _p(F6), _p(i6), _p(i2),
                                        _p(u0)), "aldic_frame_get")
        return dict(U=U, F=F, U6=U6, F6=F6, iters6=i6, iters2=i2, U0=u0)

    def frame_times(self):
        ms = np.zeros(10)
        self._ck(self.L.aldic_frame_times(self.h, _p(ms), 10), "aldic_frame_times")
        keys = ("prep", "search", "icgn6", "post6", "subpb2", "icgn2", "strain_dual", "frame",
                "search_kernel", "icgn6_kernel")
        return dict(zip(keys, ms.tolist()))

    def frame_scalars(self):
        """{beta used, ADMM steps run, RMS change of U2 and of Subpb1's U in the last
        step}."""
        v = np.zeros(4)
        self._ck(self.L.aldic_frame_scalars(self.h, _p(v), 4), "aldic_frame_scalars")
        return dict(beta=v[0], admm_steps=int(v[1]), update=(v[2], v[3]))

    def frame_counters(self):
        """{gridfit iteration budget of the next frame, frames re-run with polled
        gridfits, most gridfit iterations in the last frame, bad-node count of each ADMM
        step's Subpb1 (-1: step not run), subsets that fell back from the learned start
        (-1: the frame ran the search)}."""
        v = np.zeros(12, np.int32)
        self._ck(self.L.aldic_frame_counters(self.h, _p(v), 12), "aldic_frame_counters")
        d = dict(zip(("gf_budget", "gf_reruns", "gf_most"), v[:3].tolist()))
        d["subpb1_bad"] = v[3:11].tolist()
        d["learned_fallback"] = int(v[11])
        return d
