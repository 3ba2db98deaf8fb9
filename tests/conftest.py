import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("2d_aldic_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def speckle():
    return importlib.import_module("2d_aldic_amd.speckle")


@pytest.fixture(scope="session")
def cfg1(speckle):
    return speckle.config1()


@pytest.fixture(scope="session")
def cfg1_norm(cfg1, orc):
    f, _ = orc.normalize(cfg1.f.astype(np.float64))
    g, _ = orc.normalize(cfg1.g.astype(np.float64))
    gx, gy = orc.gradient(f)
    return f, g, gx, gy


@pytest.fixture(scope="session")
def gpu_ctx(pkg):
    return pkg.Context(0)
