import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU reference comparison")


@pytest.fixture(scope="session")
def gpu_lib():
    from hdsdp_amd.solver import load_library

    return load_library()


@pytest.fixture(scope="session")
def oracle():
    import devtools

    return devtools.oracle_lib()
