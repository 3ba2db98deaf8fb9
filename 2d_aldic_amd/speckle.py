"""Seeded synthetic speckle pairs for the AL-DIC correlation path.

Images are sums of isotropic Gaussian dots, I(X) = sum_k A_k exp(-|X - c_k|^2 / (2 s_k^2)),
evaluated in FP64.  The reference image samples I at the integer pixels; the deformed
image samples I at the inverse-mapped point X(x) of every pixel x, where x = X + u(X),
so that g(x) = f(X) holds for the analytic field u.  Both images share one scale
(full range / max of the reference), get rounded and clipped; Gaussian noise is added
to the deformed image only.

Layout: every array returned here is C-ordered with shape (ny, nx) ("row-major
[y][x]", x fastest).  That is the same memory order as the reference's MATLAB image
after its load-time transpose (dim 1 = x, ReadImage.m:79), so ``img.T`` is the
MATLAB-shaped (nx, ny) view.  Coordinates are 1-based like the reference's.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class DotField:
    cx: np.ndarray
    cy: np.ndarray
    amp: np.ndarray
    sigma: np.ndarray


def draw_dots(n_dots, nx, ny, sigma, amp=(0.4, 1.0), seed=0) -> DotField:
    """Centres uniform over [0.5, n + 0.5] first, then amplitudes, then sigmas."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0.5, nx + 0.5, n_dots)
    cy = rng.uniform(0.5, ny + 0.5, n_dots)
    a = rng.uniform(amp[0], amp[1], n_dots)
    if np.isscalar(sigma):
        s = np.full(n_dots, float(sigma))
    else:
        s = rng.uniform(sigma[0], sigma[1], n_dots)
    return DotField(cx, cy, a, s)


def _render(dots: DotField, nx, ny, Xmap=None, Ymap=None, fwd=None, chunk=8192):
    """Sum the dots on the pixel grid.

    Without maps the sample point of pixel (x, y) is (x, y) itself.  With maps it is
    (Xmap[y-1, x-1], Ymap[y-1, x-1]); ``fwd(cx, cy)`` then gives the approximate pixel
    where each dot lands so only a window around it is evaluated.
    """
    img = np.zeros(nx * ny, dtype=np.float64)
    smax = float(dots.sigma.max())
    R = int(math.ceil(5.0 * smax)) + (3 if fwd is not None else 0)
    off = np.arange(-R, R + 1)
    ox, oy = np.meshgrid(off, off, indexing="xy")
    ox = ox.ravel()
    oy = oy.ravel()
    n = dots.cx.size
    for s0 in range(0, n, chunk):
        s1 = min(n, s0 + chunk)
        cx = dots.cx[s0:s1]
        cy = dots.cy[s0:s1]
        if fwd is not None:
            px, py = fwd(cx, cy)
        else:
            px, py = cx, cy
        bx = np.rint(px).astype(np.int64)
        by = np.rint(py).astype(np.int64)
        X = bx[:, None] + ox[None, :]
        Y = by[:, None] + oy[None, :]
        ok = (X >= 1) & (X <= nx) & (Y >= 1) & (Y <= ny)
        lin = (Y - 1) * nx + (X - 1)
        lin = np.where(ok, lin, 0)
        if Xmap is None:
            sx = X.astype(np.float64)
            sy = Y.astype(np.float64)
        else:
            sx = Xmap.ravel()[lin]
            sy = Ymap.ravel()[lin]
        d2 = (sx - cx[:, None]) ** 2 + (sy - cy[:, None]) ** 2
        s = dots.sigma[s0:s1, None]
        val = dots.amp[s0:s1, None] * np.exp(-d2 / (2.0 * s * s))
        val = np.where(ok & (d2 <= (5.0 * s) ** 2), val, 0.0)
        img += np.bincount(lin.ravel(), weights=val.ravel(), minlength=nx * ny)
    return img.reshape(ny, nx)


def inverse_map(ufun, nx, ny, iters=60, tol=1e-12):
    """Solve X + u(X) = x for every pixel x by fixed-point iteration (u Lipschitz < 1)."""
    y, x = np.mgrid[1 : ny + 1, 1 : nx + 1].astype(np.float64)
    X = x.copy()
    Y = y.copy()
    for _ in range(iters):
        u, v = ufun(X, Y)
        Xn = x - u
        Yn = y - v
        err = max(np.abs(Xn - X).max(), np.abs(Yn - Y).max())
        X, Y = Xn, Yn
        if err < tol:
            break
    return X, Y


@dataclass
class SpecklePair:
    f: np.ndarray            # reference image, (ny, nx) uint8 / uint16
    g: np.ndarray            # deformed image, (ny, nx) uint8 / uint16
    ufun: object             # callable (X, Y) -> (u, v), reference coordinates
    params: dict = field(default_factory=dict)


def make_pair(nx, ny, n_dots, sigma, ufun, noise, bits=8, amp=(0.4, 1.0), seed=0):
    dots = draw_dots(n_dots, nx, ny, sigma, amp, seed)
    fr = _render(dots, nx, ny)
    full = float((1 << bits) - 1)
    scale = full / fr.max()
    Xm, Ym = inverse_map(ufun, nx, ny)

    def fwd(cx, cy):
        u, v = ufun(cx, cy)
        return cx + u, cy + v

    gd = _render(dots, nx, ny, Xm, Ym, fwd)
    rng = np.random.default_rng(seed + 1)
    dt = np.uint8 if bits <= 8 else np.uint16
    f = np.clip(np.rint(fr * scale), 0, full).astype(dt)
    gn = gd * scale
    if noise > 0:
        gn = gn + rng.normal(0.0, noise, gn.shape)
    g = np.clip(np.rint(gn), 0, full).astype(dt)
    return SpecklePair(f, g, ufun, dict(nx=nx, ny=ny, n_dots=n_dots, sigma=sigma,
                                         noise=noise, bits=bits, seed=seed))


# --- the five workloads ------------------------------------------------------

def config1(seed=0):
    """256x256 uint8, 3000 dots sigma 2, u = 2.3 + 0.001 X, v = -1.7 + 0.0005 Y, noise 1."""
    def ufun(X, Y):
        return 2.3 + 0.001 * X, -1.7 + 0.0005 * Y
    p = make_pair(256, 256, 3000, 2.0, ufun, 1.0, 8, seed=seed)
    p.params.update(ws=20, step=10, search=8, admm=3, name="config1-256-ws20-st10")
    return p


def config2(seed=0):
    """2048x2048 uint8, 150k dots sigma 2.5, u = 2 sin(2 pi X/256), v = 0.5 cos(2 pi Y/512), noise 2."""
    def ufun(X, Y):
        return 2.0 * np.sin(2 * np.pi * X / 256.0), 0.5 * np.cos(2 * np.pi * Y / 512.0)
    p = make_pair(2048, 2048, 150_000, 2.5, ufun, 2.0, 8, seed=seed)
    p.params.update(ws=32, step=16, search=8, admm=3, name="config2-2048-ws32-st16")
    return p


def config3(seed=0, n=4096, n_dots=600_000):
    """4096x4096 12-bit in uint16, 600k dots sigma U[1.5,3], 0.5% tension + Gaussian shear band."""
    c = n / 2.0
    w = 200.0
    g0 = 0.03

    def ufun(X, Y):
        from scipy.special import erf
        u = 0.005 * (X - c) + g0 * w * math.sqrt(math.pi / 2) * erf((Y - c) / (math.sqrt(2) * w))
        return u, np.zeros_like(Y)

    p = make_pair(n, n, n_dots, (1.5, 3.0), ufun, 2.0, 12, seed=seed)
    p.params.update(ws=32, step=8, search=24, admm=3, name=f"config3-{n}-ws32-st8")
    return p


def small_pair(n=96, n_dots=None, sigma=2.0, u0=1.3, v0=-0.7, ux=0.0, vy=0.0, uy=0.0, vx=0.0,
               noise=0.0, bits=8, seed=3):
    """Small affine pair for unit tests: u = u0 + ux (X-c) + uy (Y-c), v = v0 + vx (X-c) + vy (Y-c)."""
    if n_dots is None:
        n_dots = int(0.045 * n * n)
    c = (n + 1) / 2.0

    def ufun(X, Y):
        return (u0 + ux * (X - c) + uy * (Y - c), v0 + vx * (X - c) + vy * (Y - c))

    return make_pair(n, n, n_dots, sigma, ufun, noise, bits, seed=seed)
