import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: full-size configuration checks")


@pytest.fixture(scope="session")
def ipt():
    """The product package (libinplace_hip.so behind a ctypes mirror)."""
    return importlib.import_module("213transpose_amd")


@pytest.fixture(scope="session")
def oracle():
    import oracle_ctypes
    oracle_ctypes.lib()
    return oracle_ctypes
