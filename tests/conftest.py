import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE_INPUTS = "/root/reference/input_files"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def have_reference_grids():
    return os.path.isfile(os.path.join(REFERENCE_INPUTS, "grid_l3.txt"))


@pytest.fixture(scope="session")
def product():
    import geodesicodis_amd as G
    G.lib()
    return G


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


_GRIDS = {}


def generated_grid(level, lloyd=64):
    """Product-generated grid (same bits on every host) and its arrays."""
    import geodesicodis_amd as G
    key = (level, lloyd)
    if key not in _GRIDS:
        g = G.Grid.generate(level, lloyd)
        _GRIDS[key] = (g, g.arrays())
    return _GRIDS[key]


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
