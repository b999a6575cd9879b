import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP library")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    """The oracle (test infrastructure) is built from oracle/ if its library is missing."""
    lib = os.path.join(ROOT, "oracle", "build", "liborc.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    yield
