"""The device natural-neighbour refill (aldic_natural_fill: LocalICGN.m:141-161,
Subpb1Quadtree.m:188-199) against the oracle's restatement (oracle/natural_oracle.c),
FP64 on both sides: 1e-9 absolute on values of order 1.  Cases: config 1's actual
LocalICGN bad set (83 of 361 nodes, from the oracle's own IC-GN), random bad sets on
lattices with border and corner clusters (the hull rule), a quadtree mesh with hanging
nodes, scattered points, a 122 x 122 grid at 23 % bad, and the degenerate sets."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _check(gpu_ctx, orc, xy, bad, U, F, tol=1e-9):
    Ub = np.where(np.repeat(bad, 2), np.nan, U)
    Fb = np.where(np.repeat(bad, 4), np.nan, F)
    Ud, Fd = gpu_ctx.natural_fill(xy, bad, Ub, Fb)
    Uo, ko = orc.natural_fill(xy, Ub, 2, bad, return_kind=True)
    Fo = orc.natural_fill(xy, Fb, 4, bad)
    assert np.array_equal(Ud.reshape(-1, 2)[~bad], U.reshape(-1, 2)[~bad])
    fin = np.isfinite(Uo)
    assert np.array_equal(np.isfinite(Ud), fin)
    assert np.max(np.abs(Ud[fin] - Uo[fin]), initial=0.0) <= tol
    fin = np.isfinite(Fo)
    assert np.max(np.abs(Fd[fin] - Fo[fin]), initial=0.0) <= tol
    return ko


def test_config1_localicgn_bad_set(gpu_ctx, orc, cfg1_norm, speckle):
    f, g, gx, gy = cfg1_norm
    r = orc.search(f, g, [1, 256, 1, 256], 20, 10, [8, 8])
    xy = np.stack([r["x"].ravel(), r["y"].ravel()], 1)
    U0 = np.stack([r["u"].ravel(), r["v"].ravel()], 1).ravel()
    U6, F6, _, it6, _ = orc.icgn6(f, g, gx, gy, xy, U0, 20)
    bad = orc.badpoint(it6, True)
    _flush(f"config1 bad set: {bad.sum()} of {len(bad)}")
    assert bad.sum() == 83
    kind = _check(gpu_ctx, orc, xy, bad, U6, F6)
    assert (kind == 1).sum() > 40 and (kind == 2).sum() > 5


def test_lattice_random_sets(gpu_ctx, orc):
    rng = np.random.default_rng(21)
    yy, xx = np.mgrid[0:19, 0:19]
    xy = np.stack([21.0 + 10 * xx.ravel(), 21.0 + 10 * yy.ravel()], 1)
    kinds = set()
    for t in range(12):
        bad = rng.random(len(xy)) < (0.1 + 0.04 * t)
        if t % 3 == 0:
            bad[[0, 1, 19, 18, 37, 342, 360, 341]] = True   # corner clusters
        U = rng.normal(size=2 * len(xy))
        F = rng.normal(scale=0.01, size=4 * len(xy))
        _flush(f"lattice set {t}: {bad.sum()} bad")
        kinds |= set(np.unique(_check(gpu_ctx, orc, xy, bad, U, F)).tolist())
    assert {1, 2, 3} <= kinds


def test_quadtree_and_scattered(gpu_ctx, orc):
    from meshes import quadtree_mesh
    rng = np.random.default_rng(22)
    cells = [(x, y) for y in range(10) for x in range(10)]
    pick = rng.choice(len(cells), size=30, replace=False)
    xy, _ = quadtree_mesh(nc=10, h=8.0, refine=[cells[k] for k in pick])
    for t in range(4):
        bad = rng.random(len(xy)) < 0.25
        _flush(f"quadtree set {t}: {len(xy)} nodes, {bad.sum()} bad")
        _check(gpu_ctx, orc, xy, bad, rng.normal(size=2 * len(xy)),
               rng.normal(size=4 * len(xy)))
    P = rng.uniform(0, 500, (600, 2))
    for t in range(3):
        bad = rng.random(len(P)) < 0.3
        _flush(f"scattered set {t}: {bad.sum()} bad")
        _check(gpu_ctx, orc, P, bad, rng.normal(size=1200), rng.normal(size=2400))


def test_large_grid(gpu_ctx, orc):
    rng = np.random.default_rng(23)
    yy, xx = np.mgrid[0:122, 0:122]
    xy = np.stack([17.0 + 16 * xx.ravel(), 17.0 + 16 * yy.ravel()], 1)
    bad = rng.random(len(xy)) < 0.23
    bad[5000:5300] = True  # a long run of rows: large cells
    _flush(f"122 x 122 grid: {bad.sum()} bad")
    _check(gpu_ctx, orc, xy, bad, rng.normal(size=2 * len(xy)), rng.normal(size=4 * len(xy)))


def test_degenerate_sets(gpu_ctx, orc):
    xy = np.stack([np.arange(6.0), np.zeros(6)], 1)
    bad = np.array([False, True, True, False, True, False])
    kind = _check(gpu_ctx, orc, xy, bad, np.arange(12.0), np.arange(24.0))
    assert np.all(kind[bad] == 4)
    bad = np.ones(6, bool)
    Ud, Fd = gpu_ctx.natural_fill(xy, bad, np.zeros(12), np.zeros(24))
    assert np.all(np.isnan(Ud)) and np.all(np.isnan(Fd))


def test_large_bad_clusters_need_the_second_pass(gpu_ctx, orc):
    """A round hole of bad nodes in a lattice, ringed by a fine band of good nodes (a
    quadtree rim): the cells of the hole's nodes reach across it; 308 of them have more
    than 24 natural neighbours (up to 39), beyond the first pass's 24-line polygons."""
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:41, 0:41]
    xy = np.stack([10.0 + 8 * xx.ravel(), 10.0 + 8 * yy.ravel()], 1).astype(np.float64)
    c = np.array([170.0, 170.0])
    t = np.arange(120) * (2 * np.pi / 120)
    rr = 92.0 + rng.uniform(-1.5, 1.5, 120)
    rim = c + rr[:, None] * np.stack([np.cos(t), np.sin(t)], 1)
    xy = np.vstack([xy, rim])
    d = np.hypot(*(xy - c).T)
    bad = d < 88.0
    bad |= rng.random(len(xy)) < 0.03
    U = rng.normal(size=2 * len(xy))
    F = rng.normal(scale=0.01, size=4 * len(xy))
    _flush(f"hole cluster: {bad.sum()} bad of {len(xy)}")
    assert bad.sum() > 400
    kind = _check(gpu_ctx, orc, xy, bad, U, F)
    assert (kind == 1).sum() > 400
