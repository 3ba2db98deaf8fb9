"""The device frame at the benchmark workloads' full sizes and over short sequences.

* Workload 2 (2048^2 uint8, ws 32 / step 16, 14,884 subsets): the whole device frame
  against the oracle's whole frame: identical LocalICGN and Subpb1 statuses, U within
  1e-4 px and F within 1e-5.
* Workload 3 (4096^2 12-bit in uint16, shear band, ws 32 / step 8, 247,009 subsets): the
  whole device frame; against the oracle on a seeded 2 % sample of the subsets: identical
  integer-search peaks, 6-DOF IC-GN statuses equal on >= 99.9 %, RMS dU <= 0.01 px and
  |dF| <= 1e-5 where they agree; full-size properties: every subset converged, the final
  U within 0.05 px RMS of the analytic field away from the borders (the method gives ~0.023 px), and the nodal strain
  in the reference's slot order (u_y in slot 2, v_x in slot 3: the shear band's 3 % u_y
  shows up in slot 2, slot 3 stays ~0).
* Sequences (main_ALDIC.m:134-137, main_ALDIC_Quadtree.m:101-104 with NewFFTSearch = 0):
  each later frame starts from the previous frame's U; each frame against the oracle's
  frame from the same start.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_workload2_full_frame(gpu_ctx, speckle, orc):
    p = speckle.config2()
    _flush("workload 2: device frame")
    gpu_ctx.frame_stage(p.f, p.g)
    S = gpu_ctx.frame_setup(32, 16, 8, [1, 2048, 1, 2048], tol=1e-2, mu=1e-3, beta=2.56e-3,
                            n_admm=3)
    assert S == 14884
    gpu_ctx.frame_run()
    out = gpu_ctx.frame_get()
    _flush("workload 2: oracle frame")
    o = orc.frame(p.f, p.g, 32, 16, 8, [1, 2048, 1, 2048], tol=1e-2, mu=1e-3, beta=2.56e-3,
                  n_admm=3)
    assert np.array_equal(out["U0"], o["U0"])
    assert np.array_equal(out["iters6"], o["iters6"])
    assert np.array_equal(out["iters2"], o["iters2"])
    assert np.max(np.abs(out["U"] - o["U"])) < 1e-4
    assert np.max(np.abs(out["F"] - o["F"])) < 1e-5
    # the ws-32 subsets average the 256 px sinusoid (amplitude 2 px): ~0.07 px RMS
    ut, vt = p.ufun(o["xy"][:, 0], o["xy"][:, 1])
    assert np.sqrt(np.mean((out["U"][0::2] - ut) ** 2)) < 0.1


def test_workload3_full_frame(gpu_ctx, speckle, orc):
    _flush("workload 3: generating the 4096^2 pair")
    p = speckle.config3()
    gpu_ctx.frame_stage(p.f, p.g)
    S = gpu_ctx.frame_setup(32, 8, 24, [1, 4096, 1, 4096], tol=1e-2, mu=1e-3, beta=6.4e-4,
                            n_admm=3)
    assert S == 247009
    _flush("workload 3: device frame")
    gpu_ctx.frame_run()
    out = gpu_ctx.frame_get()
    it6 = out["iters6"]
    assert np.all((it6 >= 1) & (it6 <= 100))
    assert np.all(np.isfinite(out["U"])) and np.all(np.isfinite(out["F"]))
    # sampled FP64 oracle
    q = np.sort(np.random.default_rng(2026).choice(S, S // 50, replace=False))
    f, _ = orc.normalize(p.f.astype(np.float64))
    g, _ = orc.normalize(p.g.astype(np.float64))
    gx, gy = orc.gradient(f)
    _flush(f"workload 3: oracle on {q.size} subsets")
    r = orc.search_sample(f, g, [1, 4096, 1, 4096], 32, 8, [24, 24], q)
    U0o = np.stack([r["u"], r["v"]], 1)
    assert np.array_equal(out["U0"].reshape(-1, 2)[q], U0o)
    xy = np.stack([r["x"], r["y"]], 1)
    U6o, F6o, _, it6o, _ = orc.icgn6(f, g, gx, gy, xy, U0o.ravel(), 32)
    same = (it6[q] == it6o)
    assert same.mean() >= 0.999
    ok = same & (it6o >= 1) & (it6o <= 100)
    dU = out["U6"].reshape(-1, 2)[q][ok] - U6o.reshape(-1, 2)[ok]
    dF = out["F6"].reshape(-1, 4)[q][ok] - F6o.reshape(-1, 4)[ok]
    assert np.sqrt(np.mean(dU ** 2)) <= 0.01 and np.abs(dF).max() <= 1e-5
    # full-size properties of the final frame
    x1, y1, Nx, Ny = gpu_ctx.search_grid([1, 4096, 1, 4096], 32, 8, [24, 24])
    assert Nx * Ny == S
    k = np.arange(S)
    X = x1 + 8.0 * (k % Nx) + 16.0   # MeshSetUp: subset centres
    Y = y1 + 8.0 * (k // Nx) + 16.0
    _flush("workload 3: properties")
    ut, vt = p.ufun(X, Y)
    inner = (X > 80) & (X < 4016) & (Y > 80) & (Y < 4016)
    err = np.hypot(out["U"][0::2] - ut, out["U"][1::2] - vt)[inner]
    assert np.sqrt(np.mean(err ** 2)) < 0.05, np.sqrt(np.mean(err ** 2))  # method accuracy: ~0.023 px
    F = out["F"].reshape(-1, 4)
    band = inner & (np.abs(Y - 2048) < 40)
    assert np.median(F[band, 1]) > 0.02           # slot 2 holds u_y (the shear)
    assert np.abs(np.median(F[band, 2])) < 2e-3   # slot 3 holds v_x (0)
    assert abs(np.median(F[inner, 0]) - 0.005) < 1e-3  # u_x: 0.5 % tension


def test_uniform_sequence_seeded(gpu_ctx, speckle, orc):
    """Three frames of an increasing affine field on the 256^2 workload: the first from
    the integer search, the later ones from the previous frame's U."""
    prev = None
    for k, ux in enumerate((0.001, 0.002, 0.003)):
        def ufun(X, Y, ux=ux, k=k):
            return 2.3 + 0.4 * k + ux * X, -1.7 - 0.2 * k + 0.0005 * Y
        p = speckle.make_pair(256, 256, 3000, 2.0, ufun, 1.0, 8, seed=0)
        _flush(f"uniform sequence frame {k + 1}")
        gpu_ctx.frame_stage(p.f, p.g)
        gpu_ctx.frame_setup(20, 10, 8, [1, 256, 1, 256], tol=1e-2, mu=1e-3, beta=0.1, n_admm=3)
        if prev is None:
            gpu_ctx.frame_run()
        else:
            gpu_ctx.frame_run_seeded(prev)
        out = gpu_ctx.frame_get()
        o = orc.frame(p.f, p.g, 20, 10, 8, [1, 256, 1, 256], tol=1e-2, mu=1e-3, beta=0.1,
                      n_admm=3, U0=prev)
        if prev is not None:
            assert np.array_equal(out["U0"], prev)
        assert np.array_equal(out["iters6"], o["iters6"])
        assert np.array_equal(out["iters2"], o["iters2"])
        assert np.max(np.abs(out["U"] - o["U"])) < 1e-4
        ut, vt = ufun(o["xy"][:, 0], o["xy"][:, 1])
        assert np.sqrt(np.mean((out["U"][0::2] - ut) ** 2)) < 0.05
        prev = out["U"].copy()


def test_quadtree_sequence_seeded(gpu_ctx, speckle, orc):
    """Three incremental frames of the plate with a hole (far-field strain 0.4, 0.8, 1.2 %)
    on the quadtree mesh: frame 1 from the analytic field rounded to whole pixels, later
    frames from the previous frame's U; each against the oracle's quadtree frame from the
    same start."""
    from test_gpu_parity import _qframe_oracle
    qt = importlib.import_module("2d_aldic_amd.quadtree")
    n, a = 384, 60.0
    c = (n + 1) / 2.0
    xy, el, dr, ws = speckle.quadtree_hole_mesh(n, n, c, c, a, margin=48, band_near=24,
                                                band_mid=60)
    yy, xx = np.mgrid[1:n + 1, 1:n + 1]
    mask = ((xx - c) ** 2 + (yy - c) ** 2 >= a * a).astype(np.uint8)
    mu = 1e-3
    beta = 1e-3 * 16 ** 2 * mu
    prev = None
    for k, eps in enumerate((0.004, 0.008, 0.012)):
        ufun = speckle.kirsch_field(c, c, a, eps)
        p = speckle.make_pair(n, n, int(0.05 * n * n), 2.0, ufun, 1.0, 8, seed=5, hole=(c, c, a))
        _flush(f"quadtree sequence frame {k + 1}: strain {eps}")
        gpu_ctx.set_images(p.f, p.g, mask=mask)
        gpu_ctx.set_quadtree(xy, el, dr)
        if prev is None:
            u, v = ufun(xy[:, 0], xy[:, 1])
            U0 = np.column_stack([np.rint(u), np.rint(v)]).ravel()
        else:
            U0 = prev
        _, _, gxy = gpu_ctx.nodal_strain_quadtree(U0)
        tri, w = qt.linear_weights(gxy, xy)
        gpu_ctx.qframe_setup(ws, tri, w, mu=mu, beta=beta, n_admm=3, cg_rtol=1e-12)
        gpu_ctx.qframe_run(U0)
        r = gpu_ctx.qframe_get()
        f, _ = orc.normalize(p.f.astype(np.float64))
        g, _ = orc.normalize(p.g.astype(np.float64))
        gx, gy = orc.gradient(f, mask)
        o = _qframe_oracle(orc, f, g, gx, gy, xy, el, dr, ws, tri, w, U0, beta, mu, 3)
        assert np.mean(r["iters6"] == o["iters6"]) >= 0.97
        dU = np.abs(r["U"] - o["U"])
        assert np.sqrt(np.mean(dU ** 2)) <= 2e-3, (np.sqrt(np.mean(dU ** 2)), dU.max())
        far = np.hypot(xy[:, 0] - c, xy[:, 1] - c) > a + 16
        ut, vt = ufun(xy[far, 0], xy[far, 1])
        err = np.hypot(r["U"][0::2][far] - ut, r["U"][1::2][far] - vt)
        assert np.sqrt(np.mean(err ** 2)) < 0.05, (k, np.sqrt(np.mean(err ** 2)))
        prev = r["U"].copy()
    gpu_ctx.set_images(p.f, p.g)


def test_reference_quadtree_flow_plate_with_hole(gpu_ctx, speckle, orc):
    """main_ALDIC_Quadtree.m:23-33 and :64-118 on a plate with a round hole: the device
    integer search with the nodes on masked pixels emptied and the median outlier test,
    Init's in-painting, the uniform mesh refined at the hole by GenerateQuadtreeMesh (4 px
    elements at the rim, hole-edge Dirichlet nodes, unused nodes inside the hole kept) with
    U0 carried to the new nodes, then the device quadtree frame with the reference's
    uniform 32 px windows against the oracle's restatement from the same U0."""
    from test_gpu_parity import _qframe_oracle
    qt = importlib.import_module("2d_aldic_amd.quadtree")
    al = importlib.import_module("2d_aldic_amd.aldic")
    n, a = 384, 60.0
    c = (n + 1) / 2.0
    ufun = speckle.kirsch_field(c, c, a, 0.01)
    p = speckle.make_pair(n, n, int(0.05 * n * n), 2.0, ufun, 1.0, 8, seed=7, hole=(c, c, a))
    yy, xx = np.mgrid[1:n + 1, 1:n + 1]
    mask = ((xx - c) ** 2 + (yy - c) ** 2 >= a * a).astype(np.uint8)
    ws, step, s = 32, 16, 8
    _flush("reference quadtree flow: search")
    gpu_ctx.set_images(p.f, p.g, mask=mask)
    r = gpu_ctx.integer_search([1, n, 1, n], ws, step, [s, s])
    x, y, u, v = r["x"], r["y"], r["u"].copy(), r["v"].copy()
    hole = mask[y.astype(int) - 1, x.astype(int) - 1] == 0
    u[hole] = np.nan
    v[hole] = np.nan
    u, v, _ = gpu_ctx.remove_outliers(u, v, 100)
    u, v = gpu_ctx.inpaint_nans(u, 4), gpu_ctx.inpaint_nans(v, 4)
    mesh = al.MeshSetUp(x, y, dict(ImgSize=(n, n)))
    mesh["elementMinSize"] = 2
    U0u = np.column_stack([u.ravel(), v.ravel()]).ravel()
    out, U0 = qt.GenerateQuadtreeMesh(mesh, mask.T, dict(winsize=ws, winstepsize=step), U0u)
    xy = out["coordinatesFEM"]
    el = (out["elementsFEM"] - 1).astype(np.int32)
    dr = (out["dirichlet"] - 1).astype(np.int32)
    _flush(f"reference quadtree flow: {len(xy)} nodes, {len(el)} elements, {len(dr)} Dirichlet")
    gpu_ctx.set_quadtree(xy, el, dr)
    _, _, gxy = gpu_ctx.nodal_strain_quadtree(U0)
    tri, w = qt.linear_weights(gxy, xy)
    mu = 1e-3
    beta = 1e-3 * step ** 2 * mu
    wsn = np.full(len(xy), ws, np.int32)
    gpu_ctx.qframe_setup(wsn, tri, w, mu=mu, beta=beta, n_admm=3, cg_rtol=1e-12)
    gpu_ctx.qframe_run(U0)
    rq = gpu_ctx.qframe_get()
    f, _ = orc.normalize(p.f.astype(np.float64))
    g, _ = orc.normalize(p.g.astype(np.float64))
    gx, gy = orc.gradient(f, mask)
    o = _qframe_oracle(orc, f, g, gx, gy, xy, el, dr, wsn, tri, w, U0, beta, mu, 3)
    assert np.mean(rq["iters6"] == o["iters6"]) >= 0.97
    dU = np.abs(rq["U"] - o["U"])
    _flush(f"reference quadtree flow: rms dU {np.sqrt(np.mean(dU ** 2)):.2e} max {dU.max():.2e}")
    assert np.sqrt(np.mean(dU ** 2)) <= 2e-3, (np.sqrt(np.mean(dU ** 2)), dU.max())
    used = np.zeros(len(xy), bool)
    used[el[el >= 0]] = True
    far = used & (np.hypot(xy[:, 0] - c, xy[:, 1] - c) > a + 16)
    ut, vt = ufun(xy[far, 0], xy[far, 1])
    err = np.hypot(rq["U"][0::2][far] - ut, rq["U"][1::2][far] - vt)
    assert np.sqrt(np.mean(err ** 2)) < 0.05, np.sqrt(np.mean(err ** 2))
    gpu_ctx.set_images(p.f, p.g)
