"""funSmoothStrainQuadtree on the device (aldic_smooth_strain_quadtree: csrc/smooth_qt.hip)
against the CPU restatement (oracle/smooth_oracle.py: regularizeNd's normal equations by a
sparse direct solve, nanconv, two passes) on quadtree meshes with a hole and on a full
grid; FP64 preconditioned CG to a relative residual of 1e-12 against the direct solve."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))


def _field(coords, seed):
    rng = np.random.default_rng(seed)
    x, y = coords[:, 0], coords[:, 1]
    F = np.empty(4 * len(coords))
    for c in range(4):
        F[c::4] = (0.01 * (c + 1) + 1e-4 * np.sin(x / (20.0 + 5 * c)) * np.cos(y / 30.0)
                   + 2e-4 * rng.normal(size=len(coords)))
    return F


@pytest.mark.parametrize("n,a,smooth", [(256, 40.0, 1e-4), (384, 70.0, 1e-3), (512, 100.0, 1e-4)])
def test_smooth_strain_quadtree_matches_oracle(gpu_ctx, speckle, n, a, smooth):
    import smooth_oracle as so
    print(f"smooth_strain_quadtree: hole mesh n {n}, radius {a}, smoothness {smooth}", flush=True)
    c = (n + 1) / 2.0
    coords, _, _, _ = speckle.quadtree_hole_mesh(n, n, c, c, a)
    F = _field(coords, n)
    Fd, it = gpu_ctx.smooth_strain_quadtree(coords, F, 8.0, smoothness=smooth)
    Fo = so.smooth_strain_quadtree(F, coords, 8.0, smoothness=smooth)
    err = np.abs(Fd - Fo).max()
    print(f"nodes {len(coords)}, CG iterations {it.tolist()}, max |dF| {err:.2e}", flush=True)
    assert err < 1e-9
    assert np.abs(Fd - F).max() > 1e-6   # it did smooth


def test_smooth_strain_quadtree_full_grid_and_errors(gpu_ctx, pkg):
    import smooth_oracle as so
    xs = np.arange(10.0, 171.0, 8.0)
    X, Y = np.meshgrid(xs, np.arange(5.0, 126.0, 8.0), indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel()])
    F = _field(coords, 7)
    Fd, _ = gpu_ctx.smooth_strain_quadtree(coords, F, 8.0)
    Fo = so.smooth_strain_quadtree(F, coords, 8.0)
    assert np.abs(Fd - Fo).max() < 1e-9
    bad = coords.copy()
    bad[3, 1] += 0.5
    with pytest.raises(pkg.AldicError, match="off the grid"):
        gpu_ctx.smooth_strain_quadtree(bad, F, 8.0)
    with pytest.raises(pkg.AldicError, match="3 x 3"):
        gpu_ctx.smooth_strain_quadtree(coords, F, 8.0, filter_std=1.0)
