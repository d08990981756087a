import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def gpu():
    import symx_amd

    if symx_amd.device_count() < 1:
        pytest.skip("no GPU visible")
    return True


@pytest.fixture(scope="session")
def oracle():
    from oracle_py import Oracle

    return Oracle()
