import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libhr.so")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ctx():
    from vulkanhybridraytracer_amd import hr
    ctx = hr.Context(0)
    yield ctx
    ctx.close()
