import os
import sys


import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libshiro_hip.so")




@pytest.fixture(scope="session")
def hip_ctx():
    from shiro_amd.hip import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
