import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "multiq-workbench_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmqhip.so")


def pytest_collection_modifyitems(config, items):
    # keep the numpy ctypes PEP3118 warning out of the output
    import warnings
    warnings.filterwarnings("ignore", message="A builtin ctypes object gave a PEP3118")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.load()
    return oracle
