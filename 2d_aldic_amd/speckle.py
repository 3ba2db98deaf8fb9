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


_PAR = {}


def _render_group(k):
    a = _PAR
    lo, hi = a["bounds"][k], a["bounds"][k + 1]
    d = a["dots"]
    sub = DotField(d.cx[lo:hi], d.cy[lo:hi], d.amp[lo:hi], d.sigma[lo:hi])
    return _render(sub, a["nx"], a["ny"], a["Xmap"], a["Ymap"], a["fwd"])


def render_parallel(dots: DotField, nx, ny, Xmap=None, Ymap=None, fwd=None, groups=16,
                    workers=None):
    """_render over `groups` fixed, contiguous groups of dots in forked worker processes;
    the partial images are summed in group order, so the result does not depend on the
    number of workers."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import os
    n = dots.cx.size
    bounds = [n * k // groups for k in range(groups + 1)]
    _PAR.update(dots=dots, nx=nx, ny=ny, Xmap=Xmap, Ymap=Ymap, fwd=fwd, bounds=bounds)
    workers = workers or max(1, min(groups, len(os.sched_getaffinity(0)), 16))
    try:
        with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
            parts = list(ex.map(_render_group, range(groups)))
    finally:
        _PAR.clear()
    img = np.zeros((ny, nx))
    for part in parts:
        img += part
    return img


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


def make_pair(nx, ny, n_dots, sigma, ufun, noise, bits=8, amp=(0.4, 1.0), seed=0, hole=None):
    """hole = (cx, cy, a): a black disc of radius a in the reference, carried into the
    deformed image through the inverse map."""
    dots = draw_dots(n_dots, nx, ny, sigma, amp, seed)
    # large workloads render in fixed dot groups on several processes (same sum order
    # whatever the worker count); small ones in one pass
    render = render_parallel if n_dots >= 300_000 else _render
    fr = render(dots, nx, ny)
    full = float((1 << bits) - 1)
    scale = full / fr.max()
    Xm, Ym = inverse_map(ufun, nx, ny)

    def fwd(cx, cy):
        u, v = ufun(cx, cy)
        return cx + u, cy + v

    gd = render(dots, nx, ny, Xm, Ym, fwd)
    rng = np.random.default_rng(seed + 1)
    dt = np.uint8 if bits <= 8 else np.uint16
    f = np.clip(np.rint(fr * scale), 0, full).astype(dt)
    gn = gd * scale
    if noise > 0:
        gn = gn + rng.normal(0.0, noise, gn.shape)
    if hole is not None:
        hx, hy, ha = hole
        y, x = np.mgrid[1 : ny + 1, 1 : nx + 1]
        f[(x - hx) ** 2 + (y - hy) ** 2 < ha * ha] = 0
        gn[(Xm - hx) ** 2 + (Ym - hy) ** 2 < ha * ha] = 0.0
    g = np.clip(np.rint(gn), 0, full).astype(dt)
    return SpecklePair(f, g, ufun, dict(nx=nx, ny=ny, n_dots=n_dots, sigma=sigma,
                                         noise=noise, bits=bits, seed=seed))


def make_sequence(nx, ny, n_dots, sigma, ufuns, noise, bits=8, amp=(0.4, 1.0), seed=0,
                  hole=None, progress=None):
    """One reference image and a deformed image per field of ufuns (an incremental
    sequence): the dots are drawn and the reference rendered once, every deformed frame
    is rendered through its own inverse map (dot groups on several processes), with its
    own noise (generator seed + 1 + k).  progress(k), when given, is called after each
    deformed frame."""
    dots = draw_dots(n_dots, nx, ny, sigma, amp, seed)
    fr = render_parallel(dots, nx, ny)
    full = float((1 << bits) - 1)
    scale = full / fr.max()
    dt = np.uint8 if bits <= 8 else np.uint16
    f = np.clip(np.rint(fr * scale), 0, full).astype(dt)
    y, x = np.mgrid[1 : ny + 1, 1 : nx + 1]
    if hole is not None:
        hx, hy, ha = hole
        f[(x - hx) ** 2 + (y - hy) ** 2 < ha * ha] = 0
    # one frame per forked worker; each renders its dots in the same 16 groups and sums
    # them in the same order as render_parallel, so the frames are the same either way
    import concurrent.futures as cf
    import multiprocessing as mp
    import os
    n = dots.cx.size
    _SEQ.update(dots=dots, nx=nx, ny=ny, ufuns=ufuns, noise=noise, seed=seed, hole=hole,
                scale=scale, full=full, dt=dt, bounds=[n * k // 16 for k in range(17)])
    workers = max(1, min(len(ufuns), len(os.sched_getaffinity(0)), 16))
    gs = [None] * len(ufuns)
    try:
        with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
            futs = {ex.submit(_seq_frame, k): k for k in range(len(ufuns))}
            done = 0
            for fu in cf.as_completed(futs):
                gs[futs[fu]] = fu.result()
                if progress:
                    progress(done)
                done += 1
    finally:
        _SEQ.clear()
    return f, gs


_SEQ = {}


def _seq_frame(k):
    """Deformed frame k of make_sequence (run in a forked worker)."""
    a = _SEQ
    nx, ny, ufun = a["nx"], a["ny"], a["ufuns"][k]
    Xm, Ym = inverse_map(ufun, nx, ny)

    def fwd(cx, cy):
        u, v = ufun(cx, cy)
        return cx + u, cy + v

    d, b = a["dots"], a["bounds"]
    gn = np.zeros((ny, nx))
    for j in range(len(b) - 1):
        sub = DotField(d.cx[b[j]:b[j + 1]], d.cy[b[j]:b[j + 1]], d.amp[b[j]:b[j + 1]],
                       d.sigma[b[j]:b[j + 1]])
        gn += _render(sub, nx, ny, Xm, Ym, fwd)
    gn = gn * a["scale"]
    if a["noise"] > 0:
        gn = gn + np.random.default_rng(a["seed"] + 1 + k).normal(0.0, a["noise"], gn.shape)
    if a["hole"] is not None:
        hx, hy, ha = a["hole"]
        gn[(Xm - hx) ** 2 + (Ym - hy) ** 2 < ha * ha] = 0.0
    return np.clip(np.rint(gn), 0, a["full"]).astype(a["dt"])


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


def kirsch_field(cx, cy, a, eps=0.01, nu=0.3):
    """Plate with a circular hole of radius a under uniaxial tension along x, plane stress
    (Kirsch): far-field strain eps; (u, v) of the reference point (X, Y).  Inside the
    hole (masked) r is held at 0.9 a so the field stays smooth."""
    kap = (3.0 - nu) / (1.0 + nu)
    k = eps * (1.0 + nu) / 2.0  # sigma / (4 G)

    def ufun(X, Y):
        x = X - cx
        y = Y - cy
        r = np.maximum(np.hypot(x, y), 0.9 * a)
        t = np.arctan2(y, x)
        c2, s2 = np.cos(2 * t), np.sin(2 * t)
        ur = k * (r * ((kap - 1) / 2 + c2) + a * a / r * (1 + (1 + kap) * c2) - a ** 4 / r ** 3 * c2)
        ut = k * ((1 - kap) * a * a / r - r - a ** 4 / r ** 3) * s2
        return ur * np.cos(t) - ut * np.sin(t), ur * np.sin(t) + ut * np.cos(t)

    return ufun


def quadtree_hole_mesh(nx, ny, cx, cy, a, margin=64, h_far=32, h_mid=16, h_near=8,
                       band_near=64.0, band_mid=400.0):
    """Quadtree of square cells over [margin, n - margin] with a circular hole (centre
    (cx, cy), radius a): cells within band_near of the hole edge are refined to h_near,
    cells within band_mid of it to h_mid, the rest stay h_far; cells entirely inside the
    hole are dropped; neighbouring cells differ by at most one level (2:1), so an edge
    has at most one hanging node, which becomes the coarse cell's midside node.
    Returns (coords (nn, 2) 1-based pixels, elems (E, 8) 0-based with -1 for no midside
    node (corners counter-clockwise, then right, top, left, bottom midpoints), dirichlet
    = nodes inside the hole, ws = per-node window 2 x smallest incident cell clamped to
    [16, 64])."""
    n0x = (nx - 2 * margin) // h_far
    n0y = (ny - 2 * margin) // h_far
    x0 = y0 = margin

    def target(x, y, h):
        # distance range of the cell to the centre -> finest size it needs
        dx = max(cx - (x + h), 0.0, x - cx)
        dy = max(cy - (y + h), 0.0, y - cy)
        rmin = math.hypot(dx, dy)
        rmax = max(math.hypot(x - cx, y - cy), math.hypot(x + h - cx, y - cy),
                   math.hypot(x - cx, y + h - cy), math.hypot(x + h - cx, y + h - cy))
        if rmax < a:
            return -1  # inside the hole
        if rmin <= a + band_near:
            return h_near
        if rmin <= a + band_mid:
            return h_mid
        return h_far

    leaves = set()
    stack = [(x0 + i * h_far, y0 + j * h_far, h_far) for j in range(n0y) for i in range(n0x)]
    while stack:
        x, y, h = stack.pop()
        t = target(x, y, h)
        if t == -1:
            continue
        if h > t:
            hh = h // 2
            stack += [(x, y, hh), (x + hh, y, hh), (x, y + hh, hh), (x + hh, y + hh, hh)]
        else:
            leaves.add((x, y, h))

    # 2:1 balance: split a leaf while a neighbour across an edge is more than twice finer
    changed = True
    while changed:
        changed = False
        fine = {}
        for (x, y, h) in leaves:
            fine.setdefault(h, set()).add((x, y))
        split = []
        for (x, y, h) in leaves:
            if h < 4:
                continue
            # any leaf of size <= h/4 touching this cell's edges from outside?
            hit = False
            for hs in [v for v in fine if v <= h // 4]:
                for k in range(0, h, hs):
                    for (px, py) in ((x - hs, y + k), (x + h, y + k), (x + k, y - hs), (x + k, y + h)):
                        if (px, py) in fine[hs]:
                            hit = True
                            break
                    if hit:
                        break
                if hit:
                    break
            if hit:
                split.append((x, y, h))
        for (x, y, h) in split:
            leaves.discard((x, y, h))
            hh = h // 2
            for c in ((x, y), (x + hh, y), (x, y + hh), (x + hh, y + hh)):
                if target(c[0], c[1], hh) != -1:
                    leaves.add((c[0], c[1], hh))
            changed = True

    corners = set()
    for (x, y, h) in leaves:
        corners.update({(x, y), (x + h, y), (x + h, y + h), (x, y + h)})
    pts = sorted(corners, key=lambda p: (p[1], p[0]))
    index = {p: i for i, p in enumerate(pts)}
    elems = []
    cell_of = []
    for (x, y, h) in sorted(leaves, key=lambda c: (c[1], c[0], c[2])):
        hh = h // 2
        e = [index[(x, y)], index[(x + h, y)], index[(x + h, y + h)], index[(x, y + h)]]
        for m in ((x + h, y + hh), (x + hh, y + h), (x, y + hh), (x + hh, y)):
            e.append(index.get(m, -1) if hh * 2 == h else -1)
        elems.append(e)
        cell_of.append(h)
    coords = np.array(pts, dtype=np.float64)
    elems = np.array(elems, dtype=np.int64)
    hmin = np.full(len(pts), 10 ** 9)
    for e, h in zip(elems, cell_of):
        for v in e[:4]:
            hmin[v] = min(hmin[v], h)
        for v in e[4:]:
            if v >= 0:
                hmin[v] = min(hmin[v], h // 2)
    ws = np.clip(2 * hmin, 16, 64).astype(np.int32)
    r = np.hypot(coords[:, 0] - cx, coords[:, 1] - cy)
    dirichlet = np.flatnonzero(r < a)
    return coords, elems, dirichlet, ws


def config4(seed=0, n=2048, a=300.0, eps=0.01, band_near=96.0, band_mid=600.0, **mesh):
    """2048x2048 uint8 speckle (200k dots, sigma 2), circular hole radius 300 at the
    centre, Kirsch plate-with-hole field (far-field strain 1 %), noise 1; quadtree mesh
    refined towards the hole edge (cells 8-32 px, windows 16-64 px)."""
    c = (n + 1) / 2.0
    ufun = kirsch_field(c, c, a, eps)
    n_dots = int(200_000 * (n / 2048.0) ** 2)
    p = make_pair(n, n, n_dots, 2.0, ufun, 1.0, 8, seed=seed, hole=(c, c, a))
    coords, elems, dirichlet, ws = quadtree_hole_mesh(n, n, c, c, a, band_near=band_near,
                                                      band_mid=band_mid, **mesh)
    y, x = np.mgrid[1 : n + 1, 1 : n + 1]
    mask = ((x - c) ** 2 + (y - c) ** 2 >= a * a).astype(np.uint8)
    p.params.update(name=f"config4-{n}-hole{int(a)}-quadtree", coords=coords, elems=elems,
                    dirichlet=dirichlet, ws=ws, mask=mask, hole=(c, c, a))
    return p


def config4_sequence(seed=0, n=2048, a=300.0, frames=20, d_eps=0.0005, band_near=400.0,
                     band_mid=1000.0, progress=None):
    """The incremental form of workload 4: the plate with a hole at far-field strains
    d_eps, 2 d_eps, ..., frames d_eps (Kirsch fields), one reference image, and the denser
    quadtree mesh (cells 8-32 px, windows 16-64 px, ~29k nodes).  Returns (f, [g_k],
    [ufun_k], params)."""
    c = (n + 1) / 2.0
    ufuns = [kirsch_field(c, c, a, d_eps * (k + 1)) for k in range(frames)]
    n_dots = int(200_000 * (n / 2048.0) ** 2)
    f, gs = make_sequence(n, n, n_dots, 2.0, ufuns, 1.0, 8, seed=seed, hole=(c, c, a),
                          progress=progress)
    coords, elems, dirichlet, ws = quadtree_hole_mesh(n, n, c, c, a, band_near=band_near,
                                                      band_mid=band_mid)
    y, x = np.mgrid[1 : n + 1, 1 : n + 1]
    mask = ((x - c) ** 2 + (y - c) ** 2 >= a * a).astype(np.uint8)
    return f, gs, ufuns, dict(name=f"config4-seq{frames}-{n}-hole{int(a)}-quadtree", coords=coords,
                              elems=elems, dirichlet=dirichlet, ws=ws, mask=mask, hole=(c, c, a),
                              d_eps=d_eps)


def small_pair(n=96, n_dots=None, sigma=2.0, u0=1.3, v0=-0.7, ux=0.0, vy=0.0, uy=0.0, vx=0.0,
               noise=0.0, bits=8, seed=3):
    """Small affine pair for unit tests: u = u0 + ux (X-c) + uy (Y-c), v = v0 + vx (X-c) + vy (Y-c)."""
    if n_dots is None:
        n_dots = int(0.045 * n * n)
    c = (n + 1) / 2.0

    def ufun(X, Y):
        return (u0 + ux * (X - c) + uy * (Y - c), v0 + vx * (X - c) + vy * (Y - c))

    return make_pair(n, n, n_dots, sigma, ufun, noise, bits, seed=seed)
