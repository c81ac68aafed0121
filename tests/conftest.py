import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "gridapdistributed.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgdx.so on device 0)")


@pytest.fixture(scope="session")
def gpu_ctx():
    from gridap_amd import DeviceContext

    ctx = DeviceContext(0)
    yield ctx
    ctx.close()


class HostOnly:
    """Stand-in context for host-only problems (symbolic phase without a device)."""

    handle = None


@pytest.fixture
def host_ctx():
    return HostOnly()
