import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def mld():
    """The product package (HIP library through ctypes)."""
    return importlib.import_module("mld-ainv_amd")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle (checker)."""
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu(mld):
    mld.init(0)
    return mld
