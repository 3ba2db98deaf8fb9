"""CPU-side checks of the product: the HIP library loads and exports every entry
point the header declares, the host-only grid logic equals the oracle, and the host
glue (mesh set-up, inpainting) behaves like the reference."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "aldic.h")).read()
    return sorted(set(re.findall(r"\b(aldic_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pkg):
    import ctypes
    lib = pkg.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pkg._lib.EXPORTS)


def test_library_built_for_gfx950():
    so = os.path.join(ROOT, "2d_aldic_amd", "lib", "libaldic.so")
    data = open(so, "rb").read()
    assert b"gfx950" in data


def test_search_grid_host_matches_oracle(pkg, orc):
    lib = pkg.load_library()
    import ctypes as C
    for (nx, ny, grid, ws, step, s) in [(256, 256, [1, 256, 1, 256], 20, 10, [8, 8]),
                                        (2048, 2048, [1, 2048, 1, 2048], 32, 16, [8, 8]),
                                        (822, 1242, [310, 760, 70, 1146], 20, 10, [60, 60]),
                                        (300, 200, [40, 250, 30, 180], 24, 7, [5, 9])]:
        out = np.zeros(4, np.int32)
        g = np.array(grid, np.int32)
        sa = np.array(s, np.int32)
        rc = lib.aldic_search_grid(nx, ny, g.ctypes.data, ws, step, sa.ctypes.data,
                                   out.ctypes.data)
        assert rc == 0
        assert list(out) == orc.search_grid(nx, ny, grid, ws, step, s)


def test_manual_workload_subset_count(orc):
    # manual section 12 workload: ROI (309.5, 69.5)-(759.5, 1145.5), ws 20, step 10, s 60
    x1, y1, Nx, Ny = orc.search_grid(822, 1242, [310, 760, 70, 1146], 20, 10, [60, 60])
    assert (Nx, Ny) == (42, 107) and Nx * Ny == 4494


def test_meshsetup_matches_reference_layout(pkg):
    x = np.array([[11.0, 21.0, 31.0], [11.0, 21.0, 31.0]])
    y = np.array([[5.0, 5.0, 5.0], [15.0, 15.0, 15.0]])
    m = pkg.MeshSetUp(x, y, dict(winstepsize=10, ImgSize=(64, 64)))
    assert m["M"] == 3 and m["N"] == 2
    assert np.array_equal(m["coordinatesFEM"],
                          [[11, 5], [21, 5], [31, 5], [11, 15], [21, 15], [31, 15]])
    # MeshSetUp.m:37: [(j-1)M+i, (j-1)M+i+1, jM+i+1, jM+i], 1-based
    assert np.array_equal(m["elementsFEM"], [[1, 2, 5, 4], [2, 3, 6, 5]])


def test_inpaint_runs_on_the_device_only(pkg):
    # the mirror's in-painting is the library's; without a GPU it raises, never computes
    # on the host
    A = np.arange(25, dtype=float).reshape(5, 5) ** 1.5
    B = A.copy()
    B[2, 2] = np.nan
    try:
        out = pkg.inpaint_nans_springs(B)
    except pkg.AldicError:
        return
    assert out[2, 2] == pytest.approx((A[1, 2] + A[3, 2] + A[2, 1] + A[2, 3]) / 4)
    assert np.array_equal(np.delete(out.ravel(), 12), np.delete(A.ravel(), 12))


def test_no_cpu_fallback_without_gpu(pkg):
    # on a host without a GPU the context must fail loudly, never compute on the CPU
    import subprocess, sys
    code = ("import importlib,sys; sys.path.insert(0, %r); p = importlib.import_module('2d_aldic_amd');\n"
            "import os\n"
            "try:\n    p.Context(0)\nexcept p.AldicError as e:\n    print('raised')\n"
            "else:\n    print('ok')\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.stdout.strip() in ("raised", "ok")
