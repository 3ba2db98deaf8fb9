"""Quadtree mesh generation (GenerateQuadtreeMesh.m:11-150 with QrefineR.m,
provideGeometricData.m, funMarkEdge.m, funMarkInside.m), host geometry in the package's
quadtree module.  Pinned three ways:

* the reference's stored mesh: results_oht_cfrp_11_ws32_st16.mat's DICmesh (held in
  tests/golden/sample12.npz) is regenerated from its uniform grid and mask, node for node
  and element for element, with the same hole-edge (Dirichlet) list;
* a hand-derived refinement of one element of a 2 x 2 grid (new node numbers, children,
  hanging-node columns);
* mesh invariants on a plate with a round hole: every hanging node is the midpoint of
  its edge, no used node lies inside another element's edge segment, every interior
  segment is shared by exactly two elements in opposite directions, all elements keep
  the orientation, only elements of the smallest size straddle the mask boundary, and a
  linear U0 is carried to every new node exactly."""
import importlib
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sample12.npz")


@pytest.fixture(scope="module")
def qt():
    return importlib.import_module("2d_aldic_amd.quadtree")


@pytest.fixture(scope="module")
def al():
    return importlib.import_module("2d_aldic_amd.aldic")


def _uniform(al, x1, y1, Nx, Ny, ws, step, nx, ny):
    xs = x1 + ws / 2.0 + step * np.arange(Nx)
    ys = y1 + ws / 2.0 + step * np.arange(Ny)
    X, Y = np.meshgrid(xs, ys)
    mesh = al.MeshSetUp(X, Y, dict(ImgSize=(nx, ny)))
    mesh["elementMinSize"] = 2
    return mesh


def test_sample12_stored_mesh_is_regenerated(qt, al, orc):
    d = np.load(GOLD)
    ny, nx = d["f"].shape
    ws, step = int(d["winsize"]), int(d["winstepsize"])
    grid = [int(d["gridx"][0]), int(d["gridx"][1]), int(d["gridy"][0]), int(d["gridy"][1])]
    # the stored DICpara.SizeOfFFTSearchRegion is [13 13]
    x1, y1, Nx, Ny = orc.search_grid(nx, ny, grid, ws, step, [13, 13])
    mesh = _uniform(al, x1, y1, Nx, Ny, ws, step, nx, ny)
    c = d["coordinatesFEM"]
    assert np.array_equal(mesh["coordinatesFEM"], c[:Nx * Ny])  # refinement appends nodes
    out, _ = qt.GenerateQuadtreeMesh(mesh, dict(ImgRefMask=d["mask"].T),
                                     dict(winsize=ws, winstepsize=step))
    assert out["coordinatesFEM"].shape == c.shape == (1430, 2)
    assert np.array_equal(out["coordinatesFEM"], c)
    assert out["elementsFEM"].shape == (1128, 8)
    assert np.array_equal(out["elementsFEM"], d["elementsFEM"])
    assert np.array_equal(out["markCoordHoleEdge"], d["markCoordHoleEdge"])
    assert np.array_equal(out["dirichlet"], d["dirichlet"])
    assert np.array_equal(out["coordinatesFEMWorld"][:, 1], ny + 1 - c[:, 1])


def _grid_mesh(M, N, h, x0=0.0, y0=0.0):
    X, Y = np.meshgrid(x0 + h * np.arange(M), y0 + h * np.arange(N))
    xy = np.column_stack([X.ravel(), Y.ravel()])
    el = [[(j - 1) * M + i, (j - 1) * M + i + 1, j * M + i + 1, j * M + i]
          for j in range(1, N) for i in range(1, M)]
    return xy, np.array(el, np.int64)


def test_qrefine_one_element_known_answer(qt):
    # nodes 1..9 (x fastest, spacing 8); elements e1 = (1 2 5 4), e2 = (2 3 6 5),
    # e3 = (4 5 8 7), e4 = (5 6 9 8); refine e1.  Sorted unique edges: (1,2) (1,4) (2,3)
    # (2,5) (3,6) (4,5) (4,7) (5,6) (5,8) (6,9) (7,8) (8,9); e1's edges (1,2) (1,4) (2,5)
    # (4,5) get nodes 10 11 12 13 in that order, the centroid is 14.
    xy, el = _grid_mesh(3, 3, 8.0)
    c, e, irr = qt.QrefineR(xy, el, np.zeros((0, 3)), [1])
    assert c.shape == (14, 2)
    assert np.array_equal(c[9:], [[4, 0], [0, 4], [8, 4], [4, 8], [4, 4]])
    assert np.array_equal(e, [[1, 10, 14, 11], [2, 12, 14, 10], [5, 13, 14, 12],
                              [4, 11, 14, 13], [2, 3, 6, 5], [4, 5, 8, 7], [5, 6, 9, 8]])
    # hanging nodes in MATLAB find order over the unrefined elements with a new edge node:
    # column 1 first (e3's edge (4,5) -> 13), then column 4 (e2's edge (5,2) -> 12)
    assert np.array_equal(irr, [[4, 5, 13], [2, 5, 12]])
    e8 = qt._hanging_columns(e, irr)
    assert np.array_equal(e8[4], [2, 3, 6, 5, 0, 0, 12, 0])   # edge (4,1) = (5,2): column 7
    assert np.array_equal(e8[5], [4, 5, 8, 7, 0, 0, 0, 13])   # edge (1,2) = (4,5): column 8
    assert not e8[[0, 1, 2, 3, 6], 4:].any()


def test_qrefine_closure_forces_coarse_neighbours(qt):
    # after e1 is split, refining its child (5 13 14 12) would put a second level on the
    # hanging edges (4,5) and (2,5): e3 and e2 must be split with it, e4 keeps two hanging
    # nodes (one from each) and nothing else changes
    xy, el = _grid_mesh(3, 3, 8.0)
    c, e, irr = qt.QrefineR(xy, el, np.zeros((0, 3)), [1])
    c2, e2, irr2 = qt.QrefineR(c, e, irr, [3])
    assert e2.shape[0] == 3 + 4 + 4 + 4 + 1
    assert np.array_equal(e2[:2], e[:2]) and np.array_equal(e2[6], e[3])
    assert np.array_equal(e2[-1], [5, 6, 9, 8])
    e8 = qt._hanging_columns(e2, irr2)
    assert np.count_nonzero(e8[-1, 4:]) == 2
    _check_mesh(c2, e8)


def _segments(xy, e8):
    """Directed boundary segments of each element: an edge, or its two halves when it
    holds a hanging node (columns 8, 5, 6, 7 for edges (1,2), (2,3), (3,4), (4,1))."""
    seg = []
    for row in e8:
        for k, col in enumerate((7, 4, 5, 6)):
            a, b, h = row[k], row[(k + 1) % 4], row[col]
            if h:
                assert np.allclose(xy[h - 1], 0.5 * (xy[a - 1] + xy[b - 1]))
                seg += [(a, h), (h, b)]
            else:
                seg.append((a, b))
    return seg


def _check_mesh(xy, e8, outer=None):
    seg = _segments(xy, e8)
    count = {}
    for a, b in seg:
        count.setdefault((min(a, b), max(a, b)), []).append((a, b))
    for k, v in count.items():
        assert len(v) <= 2, k
        if len(v) == 2:
            assert v[0] == (v[1][1], v[1][0]), v        # opposite directions
    used = np.unique(e8[e8 > 0])
    P = xy[used - 1]
    for (a, b) in count:                                  # no node strictly inside a segment
        A, B = xy[a - 1], xy[b - 1]
        AB = B - A
        t = (P - A) @ AB / (AB @ AB)
        d = np.abs((P[:, 0] - A[0]) * AB[1] - (P[:, 1] - A[1]) * AB[0])
        assert not np.any((d < 1e-9) & (t > 1e-9) & (t < 1 - 1e-9)), (a, b)
    q = xy[e8[:, :4] - 1]                                 # shoelace: one orientation
    area = 0.5 * ((q[:, :, 0] * np.roll(q[:, :, 1], -1, 1)).sum(1)
                  - (q[:, :, 1] * np.roll(q[:, :, 0], -1, 1)).sum(1))
    assert np.all(area > 0) or np.all(area < 0)
    return count


def test_plate_with_hole_mesh_invariants(qt, al):
    nx = ny = 256
    c, a = 128.5, 50.0
    yy, xx = np.mgrid[1:ny + 1, 1:nx + 1]
    mask = ((xx - c) ** 2 + (yy - c) ** 2 >= a * a).astype(np.float64)
    ws, step = 32, 16
    mesh = _uniform(al, 16, 16, 13, 13, ws, step, nx, ny)       # nodes 32, 48, ..., 224
    u0 = lambda x, y: 0.3 + 0.01 * x - 0.02 * y
    v0 = lambda x, y: -1.1 + 0.005 * x + 0.015 * y
    xy0 = mesh["coordinatesFEM"]
    U0 = np.column_stack([u0(xy0[:, 0], xy0[:, 1]), v0(xy0[:, 0], xy0[:, 1])]).ravel()
    out, U = qt.GenerateQuadtreeMesh(mesh, mask.T, dict(winsize=ws, winstepsize=step), U0)
    xy, e8 = out["coordinatesFEM"], out["elementsFEM"]
    assert e8.shape[1] == 8 and np.count_nonzero(e8[:, 4:]) > 0
    count = _check_mesh(xy, e8)
    q = xy[e8[:, :4] - 1]
    assert set(np.abs(q[:, 0, 0] - q[:, 2, 0]).tolist()) == {4.0, 8.0, 16.0}
    # the elements that straddle the hole edge are all of the smallest size (4 px)
    x0i, x1i = q[:, :, 0].min(1).astype(int), q[:, :, 0].max(1).astype(int)
    y0i, y1i = q[:, :, 1].min(1).astype(int), q[:, :, 1].max(1).astype(int)
    for k in range(e8.shape[0]):
        box = mask[y0i[k] - 1:y1i[k], x0i[k] - 1:x1i[k]]
        if box.min() != box.max():
            assert x1i[k] - x0i[k] <= 4 and y1i[k] - y0i[k] <= 4
    # every remaining element is mostly outside the hole; the removed ones are inside
    assert len(np.unique(e8[e8 > 0])) < xy.shape[0]               # hole nodes kept, unused
    # once-used segments are on the outer border or on the hole's rim
    for (p, r), v in count.items():
        if len(v) == 1:
            m = 0.5 * (xy[p - 1] + xy[r - 1])
            on_border = (m[0] in (32, 224)) or (m[1] in (32, 224))
            assert on_border or abs(np.hypot(m[0] - c, m[1] - c) - a) < 8, m
    # the hole-edge set holds every node of the refined elements and is closed twice
    small = np.hypot(*(q[:, 0] - q[:, 2]).T) < 0.99 * np.sqrt(2) * step
    assert np.all(np.isin(e8[small][e8[small] > 0], out["markCoordHoleEdge"]))
    assert np.array_equal(out["dirichlet"], out["markCoordHoleEdge"])
    # a linear field is carried to every node exactly
    assert np.allclose(U[0::2], u0(xy[:, 0], xy[:, 1]), atol=1e-12)
    assert np.allclose(U[1::2], v0(xy[:, 0], xy[:, 1]), atol=1e-12)


def test_mark_edge_and_inside_rules(qt):
    # one 8 x 8 element over a mask with a single zero pixel: marked (range > 0, size 8 >
    # 2 * 2) and outside (sum 80 of 81 pixels >= 0.5 * 64); with a 4 px element, not marked;
    # a fully masked element is inside
    xy = np.array([[1, 1], [9, 1], [9, 9], [1, 9]], np.float64)
    el = np.array([[1, 2, 3, 4]])
    m = np.ones((9, 9))
    m[4, 4] = 0
    assert qt.funMarkEdge(xy, el, m, 4)[1].tolist() == [True]
    assert qt.funMarkEdge(xy, el, m, 8)[1].tolist() == [False]
    assert qt.funMarkInside(xy, el, m)[1].tolist() == [1]
    assert qt.funMarkInside(xy, el, np.zeros((9, 9)))[0].tolist() == [1]
    # the threshold compares the sum over the (w+1)^2 pixel box with half of w^2
    m2 = np.zeros((9, 9))
    m2[:, :4] = 1                                         # 36 ones, 0.5 * 64 = 32
    assert qt.funMarkInside(xy, el, m2)[1].tolist() == [1]
    m2[:, 3] = 0                                          # 27 ones
    assert qt.funMarkInside(xy, el, m2)[0].tolist() == [1]
