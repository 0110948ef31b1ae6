import os
import sys

import pytest

# torch ships its own HIP runtime: loaded after librome_gpu.so's (/opt/rocm), torch finds no GPU; loaded
# first, both work.  Tests that put arrays on the device with torch rely on this order.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def engine():
    import rome_amd
    e = rome_amd.Engine(0)
    yield e
    e.close()
