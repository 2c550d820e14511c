import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__ as G  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs the libpbf.so kernels)")


@pytest.fixture(scope="session")
def pkg():
    return G.load_package()


@pytest.fixture(scope="session")
def oracle():
    O = G.load_oracle()
    O.lib()
    return O


@pytest.fixture(scope="session")
def scene(pkg):
    return pkg.scene


@pytest.fixture(scope="session")
def fluid(pkg):
    from importlib import import_module
    return import_module(G.PKG_NAME + ".fluid")


@pytest.fixture(scope="session")
def capi(pkg):
    from importlib import import_module
    return import_module(G.PKG_NAME + "._capi")
