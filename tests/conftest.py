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


H5DUMP = "/opt/conda/bin/h5dump"


def h5_read(path, name):
    """A float32 dataset of an HDF5 file, through h5dump's text output with
    9 significant digits (enough to round-trip float32; no Python HDF5 module
    is installed)."""
    import re
    import subprocess
    out = subprocess.run([H5DUMP, "-m", "%.9g", "-d", name, str(path)], capture_output=True, text=True,
                         check=True).stdout
    dims = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", out)
    shape = tuple(int(x) for x in dims.group(1).split(","))
    body = out[out.index("DATA {") + len("DATA {"):out.rindex("}")]
    vals = []
    for line in body.splitlines():
        line = re.sub(r"^\s*\([^)]*\):", "", line).strip().rstrip("}").strip()
        vals.extend(float(v) for v in line.split(",") if v.strip())
    return np.array(vals, np.float64).astype(np.float32).reshape(shape)
