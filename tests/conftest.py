import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (REPO, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU runs (tens of seconds)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


@pytest.fixture(scope="session")
def gauss_table():
    return load_golden("gauss_ref")


def host_for(g):
    """tritok_amd Host with the mesh, boundary conditions and tri.inp of a fixture."""
    from tritok_amd import Host
    h = Host.from_arrays(g["mesh_x"], g["mesh_ibc"], g["mesh_idc"], g["mesh_bewn"], int(g["mesh_nbnodes"]))
    h.read_bic(g["bic_text"].tobytes().decode()).read_inp(g["inp_text"].tobytes().decode())
    return h


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / nb if nb > 0 else np.linalg.norm(a)
