import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bandedmatrices.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libbandedhip.so")


@pytest.fixture(scope="session")
def bh():
    """The product package with its library loaded and the device initialised."""
    import banded_hip

    L = banded_hip.load_library()
    st = L.bh_init(0)
    assert st == 0, L.bh_last_error().decode()
    return banded_hip


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O
