"""Pinning of the oracle's natural-neighbour refill (oracle/natural_oracle.c), the
reference's scatteredInterpolant(..., 'natural') of LocalICGN.m:141-161 and
Subpb1Quadtree.m:188-199.  None of these checks restates the clipping code: they use
closed forms (an isolated lattice hole takes 1/4 from each edge neighbour; the hull rule's
parallelogram and edge-interpolation forms), invariants (weights >= 0, sum 1, linear
precision, so linear fields come back exactly) and a brute-force sampling of the
nearest-site map (the definition of Sibson's coordinates as stolen Voronoi areas)."""
import numpy as np
import pytest


def lattice(nx, ny, h=1.0, x0=0.0, y0=0.0):
    yy, xx = np.mgrid[0:ny, 0:nx]
    return np.stack([x0 + h * xx.ravel(), y0 + h * yy.ravel()], 1).astype(np.float64)


def _sample_cell(P, q, R, res):
    t = (np.arange(res) + 0.5) / res * 2 * R - R
    X, Y = np.meshgrid(q[0] + t, q[1] + t)
    xs = np.stack([X.ravel(), Y.ravel()], 1)
    w = np.zeros(len(P))
    ext = 0.0
    for c in range(0, xs.shape[0], 200000):
        x = xs[c:c + 200000]
        dq = ((x - q) ** 2).sum(1)
        dp = ((x[:, None, :] - P[None, :, :]) ** 2).sum(2)
        j = dp.argmin(1)
        inside = dq < dp[np.arange(len(x)), j]
        w += np.bincount(j[inside], minlength=len(P))
        if inside.any():
            ext = max(ext, np.abs(x[inside] - q).max())
    return w, ext


def brute_sibson(P, q, res=1400):
    """Stolen-area fractions by sampling: a sample x belongs to the cell of q when q is
    nearer than every site of P; it is stolen from its nearest site in P.  A coarse pass
    finds the cell's extent, the fine pass samples a square just around it."""
    d = np.sqrt(((P - q) ** 2).sum(1))
    R0 = 3.0 * np.sort(d)[min(8, len(d) - 1)]
    _, ext = _sample_cell(P, q, R0, 300)
    assert ext < 0.9 * R0, "cell reaches the coarse sampling square"
    w, ext2 = _sample_cell(P, q, 1.05 * ext + 2 * R0 / 300, res)
    return w / w.sum()


def test_isolated_hole_quarter_weights(orc):
    for h in (1.0, 10.0, 16.0):
        P = lattice(7, 7, h, 3.0, 5.0)
        q = P[24].copy()
        keep = np.ones(len(P), bool)
        keep[24] = False
        idx, w = orc.sibson(P[keep], q)
        full = np.zeros(len(P))
        full[np.flatnonzero(keep)[idx]] = w
        for nb in (17, 23, 25, 31):  # the four edge neighbours
            assert abs(full[nb] - 0.25) < 1e-13
        assert abs(full.sum() - 1.0) < 1e-13
        assert np.all(full >= -1e-15)


def test_weights_partition_and_linear_precision(orc):
    rng = np.random.default_rng(3)
    for trial in range(30):
        P = rng.uniform(0, 10, (40, 2))
        q = rng.uniform(4, 6, 2)
        idx, w = orc.sibson(P, q)
        assert np.all(w >= -1e-14)
        assert abs(w.sum() - 1.0) < 1e-12
        assert np.allclose((w[:, None] * P[idx]).sum(0), q, atol=1e-10)


def test_matches_brute_force_area_sampling(orc):
    rng = np.random.default_rng(11)
    cases = [rng.uniform(0, 10, (25, 2)) for _ in range(3)]
    L = lattice(6, 6)
    keep = np.ones(36, bool)
    keep[[14, 15, 21]] = False  # L-shaped cluster; q = node 14
    cases.append((L[keep], L[14]))
    for c in cases:
        if isinstance(c, tuple):
            P, q = c
        else:
            P, q = c, np.array([5.1, 4.7])
        idx, w = orc.sibson(P, q)
        full = np.zeros(len(P))
        full[idx] = w
        ref = brute_sibson(P, q)
        assert np.max(np.abs(full - ref)) < 1e-3, (full, ref)


@pytest.mark.parametrize("cluster", [
    [24],                      # isolated
    [24, 25],                  # pair
    [24, 25, 32],              # L-shaped
    [16, 17, 18, 23, 24, 25, 30, 31, 32],  # 3 x 3
])
def test_linear_fields_reproduced_inside(orc, cluster):
    P = lattice(7, 7, 10.0, 21.0, 21.0)
    bad = np.zeros(len(P), bool)
    bad[cluster] = True
    V = np.stack([0.3 + 0.01 * P[:, 0] - 0.02 * P[:, 1], -1.7 + 0.005 * P[:, 0] + 0.003 * P[:, 1]],
                 1).ravel()
    Vin = V.copy()
    Vin.reshape(-1, 2)[bad] = np.nan
    out, kind = orc.natural_fill(P, Vin, 2, bad, return_kind=True)
    assert np.all(kind[bad] == 1) and np.all(kind[~bad] == 0)
    assert np.max(np.abs(out - V)) < 1e-12


def test_random_clusters_and_hull_cases_linear_exact(orc):
    rng = np.random.default_rng(5)
    P = lattice(19, 19, 10.0, 21.0, 21.0)
    for trial in range(20):
        bad = rng.random(len(P)) < 0.25
        bad[[0, 18, 342, 360]] = rng.random(4) < 0.5  # corners sometimes bad
        V = np.stack([1.0 + 0.01 * P[:, 0] + 0.002 * P[:, 1], 2.0 - 0.004 * P[:, 0]], 1).ravel()
        out, kind = orc.natural_fill(P, np.where(np.repeat(bad, 2), np.nan, V), 2, bad,
                                     return_kind=True)
        assert np.max(np.abs(out - V)) < 1e-11
        assert set(np.unique(kind[bad])) <= {1, 2, 3}


def test_hull_rule_closed_forms(orc):
    P = lattice(6, 5)
    f = lambda x, y: np.sin(x) + y * y  # nonlinear: pins the rule, not linear precision
    V = f(P[:, 0], P[:, 1])
    # corner (0, 0) bad: nearest hull point is inside the edge (1,0)-(0,1); the Delaunay
    # third vertex is (1,1) -> parallelogram rule
    bad = np.zeros(len(P), bool)
    bad[0] = True
    out, kind = orc.natural_fill(P, np.where(bad, np.nan, V), 1, bad, return_kind=True)
    assert kind[0] == 2
    assert abs(out[0] - (f(1, 0) + f(0, 1) - f(1, 1))) < 1e-13
    # bottom-border run (2,0), (3,0) bad: on the hull boundary -> linear along the edge
    bad = np.zeros(len(P), bool)
    bad[[2, 3]] = True
    out, kind = orc.natural_fill(P, np.where(bad, np.nan, V), 1, bad, return_kind=True)
    assert kind[2] == 2 and kind[3] == 2
    assert abs(out[2] - (2 * f(1, 0) + f(4, 0)) / 3) < 1e-13
    assert abs(out[3] - (f(1, 0) + 2 * f(4, 0)) / 3) < 1e-13
    # left column and bottom row bad: the good nodes' hull is [1,5] x [1,4] and q = (0, 0)
    # lies beyond its vertex (1, 1) -> plane through (1,1) and its hull neighbours (2,1)
    # and (1,2)
    bad = np.zeros(len(P), bool)
    for k in range(len(P)):
        x, y = P[k]
        if (x == 0 or y == 0):
            bad[k] = True
    out, kind = orc.natural_fill(P, np.where(bad, np.nan, V), 1, bad, return_kind=True)
    assert kind[0] == 3
    want = f(1, 1) + (f(2, 1) - f(1, 1)) * (0 - 1) + (f(1, 2) - f(1, 1)) * (0 - 1)
    assert abs(out[0] - want) < 1e-13


def test_degenerate_sets(orc):
    P = lattice(4, 1)
    bad = np.array([False, True, True, False])
    V = np.array([1.0, 0, 0, 4.0])
    out, kind = orc.natural_fill(P, V, 1, bad, return_kind=True)
    assert np.all(kind[bad] == 4)
    assert out[1] == 1.0 and out[2] == 4.0  # nearest good node
    out, kind = orc.natural_fill(P, V, 1, np.ones(4, bool), return_kind=True)
    assert np.all(np.isnan(out)) and np.all(kind == 5)
