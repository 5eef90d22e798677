import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def _read_vtk(path):
    """ASCII VTK legacy unstructured grid (the layout base::io::vtk::
    LegacyWriter writes, base/io/vtk/LegacyWriter.hpp:150-296): points
    (n, 3), cells (list of id lists), cell types, and the point / cell data
    arrays by name, one row per point / cell."""
    with open(path) as fh:
        tok = fh.read().split("\n")
    out = {"point_data": {}, "cell_data": {}}
    i, section = 0, None
    while i < len(tok):
        ln = tok[i].split()
        i += 1
        if not ln:
            continue
        key = ln[0]
        if key == "POINTS":
            n = int(ln[1])
            out["points"] = np.array([[float(v) for v in tok[i + k].split()] for k in range(n)])
            i += n
        elif key == "CELLS":
            n = int(ln[1])
            out["cells"] = [[int(v) for v in tok[i + k].split()[1:]] for k in range(n)]
            i += n
        elif key == "CELL_TYPES":
            n = int(ln[1])
            out["cell_types"] = [int(tok[i + k]) for k in range(n)]
            i += n
        elif key in ("POINT_DATA", "CELL_DATA"):
            section = ("point_data" if key == "POINT_DATA" else "cell_data", int(ln[1]))
        elif key in ("SCALARS", "VECTORS", "TENSORS"):
            if key == "SCALARS":
                i += 1   # LOOKUP_TABLE line
            name, n = ln[1], section[1]
            out[section[0]][name] = np.array([[float(v) for v in tok[i + k].split()]
                                              for k in range(n)])
            i += n
    return out


@pytest.fixture(scope="session")
def read_vtk():
    return _read_vtk
