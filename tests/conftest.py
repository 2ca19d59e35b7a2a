import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through libfcw.so)")


@pytest.fixture(scope="session")
def surf_c1():
    from fastcorticalwiring_amd import meshes

    return meshes.config_surface(0)


@pytest.fixture(scope="session")
def surf_c2():
    from fastcorticalwiring_amd import meshes

    return meshes.config_surface(1)


def rel_err(a, b, floor=1e-12):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)
