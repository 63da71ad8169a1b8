import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")


@pytest.fixture(scope="session")
def vsgm():
    import vsgm_amd

    vsgm_amd.api.lib()
    return vsgm_amd


@pytest.fixture(scope="session")
def oracle():
    import oracle

    return oracle


@pytest.fixture(scope="session")
def gpu(vsgm):
    n = vsgm.device_count()
    assert n > 0, "GPU tests need a visible MI355X"
    print(f"[gpu] {n} device(s) visible", flush=True)
    return n
