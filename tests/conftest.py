import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (an MI355X)")
    config.addinivalue_line("markers", "slow: large inputs (seconds of CPU work)")


@pytest.fixture(scope="session")
def msc():
    import connectivity_amd

    return connectivity_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx(msc):
    if msc.device_count() < 1:
        pytest.fail("no gfx950 device visible: GPU tests need an MI355X")
    ctx = msc.Context(0)
    yield ctx
    ctx.close()
