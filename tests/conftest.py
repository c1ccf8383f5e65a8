import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def fixture_geometry():
    import _oracle
    return _oracle.load_fixture_geometry()


@pytest.fixture(scope="session")
def gold():
    import numpy as np
    import _oracle
    out = {}
    for kind in ("transport", "adsorption"):
        for name in ("ux", "uy", "rho"):
            out[(kind, name)] = np.loadtxt(os.path.join(_oracle.FIXTURES, "gold", kind, name + ".txt"))
    return out
