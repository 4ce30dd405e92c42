import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _gpu_visible() -> bool:
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def device():
    from birp_cpp_amd import capi
    if not _gpu_visible():
        pytest.skip("no GPU visible")
    dev = capi.Device(0)
    yield dev
    dev.close()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    return O
