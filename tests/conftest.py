import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built librtq.so")


@pytest.fixture(scope="session", autouse=True)
def _host_libraries():
    """The host builder and the oracle are plain C/C++; build them here if they are missing."""
    from tim_raytracing_amd import build as nb
    if not os.path.exists(os.path.join(ROOT, "tim_raytracing_amd", "librtq_build.so")):
        nb.build_host()
    if not os.path.exists(os.path.join(ROOT, "oracle", "liborc.so")):
        nb.build_oracle()
    yield
