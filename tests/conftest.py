import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "compute.scala_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_tensors():
    """The back-end on GPU 0. Fails (does not skip) when the library or the GPU is missing."""
    from compute import Tensors
    t = Tensors(0)
    yield t
    t.close()


@pytest.fixture(scope="session")
def offline_tensors():
    from compute import Tensors
    return Tensors(offline=True)
