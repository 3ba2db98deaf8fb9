"""Strain post-processing and the sequence driver on the device against the oracle:
aldic_compute_strain (ComputeStrain.m's four methods and three strain types, the
trimmed world output), the host mirror ComputeStrain, and run_sequence's result
containers (main_ALDIC.m:40-43,455-457)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(M, N, h, rng):
    i, j = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    x, y = h * i, h * j
    u = 0.5 + 0.01 * x - 0.004 * y + 2e-4 * x * y / h + rng.normal(scale=0.02, size=x.shape)
    v = -0.3 + 0.006 * x + 0.02 * y - 1e-3 * np.sin(x / 17.0) + rng.normal(scale=0.02, size=x.shape)
    U = np.empty(2 * M * N)
    U[0::2] = u.flatten(order="F")
    U[1::2] = v.flatten(order="F")
    return U


def _q4(M, N, h):
    i, j = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    coords = np.stack([h * i.flatten(order="F") + 11.0, h * j.flatten(order="F") + 7.0], 1)
    el = []
    for jj in range(N - 1):
        for ii in range(M - 1):
            n = ii + M * jj
            el.append([n, n + 1, n + 1 + M, n + M])
    return coords, np.array(el, np.int32)


@pytest.mark.parametrize("method,rad", [(0, 0), (1, 1), (2, 1), (2, 3), (3, 1)])
@pytest.mark.parametrize("stype", [0, 1, 2])
def test_compute_strain_matches_oracle(gpu_ctx, orc, method, rad, stype):
    rng = np.random.default_rng(10 * method + rad + 100 * stype)
    M, N, h = 23, 17, 8.0
    U = _field(M, N, h, rng)
    F = rng.normal(scale=0.01, size=4 * M * N)
    coords, el = _q4(M, N, h)
    if method == 3:
        gpu_ctx.set_mesh(coords, el)
    Ff, Fw, r = gpu_ctx.compute_strain(M, N, h, U=U, F=F, method=method, rad=rad,
                                       strain_type=stype)
    o = orc.compute_strain(M, N, h, U=U, F=F, method=method, rad=rad, stype=stype,
                           coords=coords, elems=el)
    assert r == o["Rad"]
    scale = np.max(np.abs(o["FStrainFinite"]))
    assert np.max(np.abs(Ff - o["FStrainFinite"])) <= 1e-12 * scale
    assert Fw.shape == o["FStrainWorld"].shape
    assert np.max(np.abs(Fw - o["FStrainWorld"])) <= 1e-12 * scale


def test_compute_strain_edge_cases(gpu_ctx, orc):
    rng = np.random.default_rng(7)
    # one node line in x: no x derivative (the operator has no entries there)
    U = rng.normal(size=2 * 1 * 6)
    Ff, Fw, r = gpu_ctx.compute_strain(1, 6, 3.0, U=U, method=1)
    o = orc.compute_strain(1, 6, 3.0, U=U, method=1)
    assert np.allclose(Ff, o["FStrainFinite"], atol=1e-14) and Fw.size == 0 and r == 1
    assert np.all(Ff[0::4] == 0) and np.all(Ff[1::4] == 0)
    # plane fit wider than the grid allows, unknown strain type, method 3 without a mesh
    from importlib import import_module
    err = import_module("2d_aldic_amd").AldicError
    with pytest.raises(err):
        gpu_ctx.compute_strain(5, 5, 1.0, U=rng.normal(size=50), method=2, rad=3)
    with pytest.raises(err):
        gpu_ctx.compute_strain(5, 5, 1.0, U=rng.normal(size=50), method=1, strain_type=3)


def test_compute_strain_mirror(pkg, orc):
    rng = np.random.default_rng(11)
    M, N, h = 12, 9, 5.0
    U = _field(M, N, h, rng)
    out = pkg.ComputeStrain(dict(MethodToComputeStrain=2, StrainType=2, winstepsize=h), U,
                            None, M, N, Rad=2)
    o = orc.compute_strain(M, N, h, U=U, method=2, rad=2, stype=2)
    assert out["Rad"] == 2 and np.array_equal(out["temp3"], o["temp3"])
    assert np.allclose(out["FStraintemp"], out["FStrainFinite"][out["temp3"]], atol=0)
    assert np.allclose(out["FStrainWorld"], o["FStrainWorld"], atol=1e-14)


def test_run_sequence_result_containers(pkg, speckle, orc):
    """Three deformed frames of an increasing affine field (cumulative mode): the first
    from the integer search, the later ones seeded with the previous U; the containers
    hold what main_ALDIC.m stores, and the seeded frame equals the oracle's frame."""
    frames = []
    fields = []
    for k, ux in enumerate((0.001, 0.002, 0.003)):
        def ufun(X, Y, ux=ux, k=k):
            return 2.3 + 0.4 * k + ux * X, -1.7 - 0.2 * k + 0.0005 * Y
        p = speckle.make_pair(256, 256, 3000, 2.0, ufun, 1.0, 8, seed=0)
        if k == 0:
            frames.append(p.f)
        frames.append(p.g)
        fields.append(ufun)
    DICpara = dict(winsize=20, winstepsize=10, SizeOfFFTSearchRegion=8, tol=1e-2, mu=1e-3,
                   beta=0.1, ALSolveStep=3)
    res = pkg.run_sequence(frames, DICpara)
    RD, RF, RM, RE = (res[k] for k in ("ResultDisp", "ResultDefGrad", "ResultFEMesh",
                                       "ResultFEMeshEachFrame"))
    S = res["DICmesh"]["coordinatesFEM"].shape[0]
    assert len(RD) == len(RF) == len(RE) == 3 and len(RM) == 1
    assert set(RM[0]) == {"coordinatesFEM", "elementsFEM", "winsize", "winstepsize",
                          "gridxyROIRange"}
    assert RM[0]["elementsFEM"].min() == 1 and RM[0]["elementsFEM"].shape[1] == 4
    for k in range(3):
        assert RD[k]["U"].shape == (2 * S,) and RF[k]["F"].shape == (4 * S,)
        assert RD[k]["ALSub1BadPtNum"].shape == (6,)
        xy = res["DICmesh"]["coordinatesFEM"]
        ut, vt = fields[k](xy[:, 0], xy[:, 1])
        assert np.sqrt(np.mean((RD[k]["U"][0::2] - ut) ** 2)) < 0.05
    # frame 2 from frame 1's U, against the oracle's frame with the same initial guess
    o = orc.frame(frames[0], frames[2], 20, 10, 8, [1, 256, 1, 256], tol=1e-2, mu=1e-3,
                  beta=0.1, n_admm=3, U0=RD[0]["U"])
    assert np.max(np.abs(RD[1]["U"] - o["U"])) < 1e-4
    it6, it2 = o["iters6"], o["iters2"]
    assert RD[1]["ALSub1BadPtNum"][0] == np.sum((it6 < 0) | (it6 > 99)) - np.sum(it6 == 102)
    assert RD[1]["ALSub1BadPtNum"][3] == np.sum((it2 < 0) | (it2 > 99))
    assert np.all(RD[1]["ALSub1BadPtNum"][4:] == 0)
    # the coordinates are the search grid's subset centres, x fastest
    xy = res["DICmesh"]["coordinatesFEM"]
    assert np.allclose(xy[:, 0].reshape(-1, o["Nx"])[0], o["xy"][:o["Nx"], 0])
    assert np.allclose(xy, o["xy"])
