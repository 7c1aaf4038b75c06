import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def uepm():
    """The product library through its C ABI (ultimateepm_amd.api)."""
    from ultimateepm_amd import api

    return api


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    from ultimateepm_amd import api

    ctx = api.Context(0)
    yield ctx
    ctx.close()
