import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch (used by GPU tests for device buffers) carries its own HIP runtime; it
# must be loaded before libgpzip_hip.so pulls the system one into the process,
# exactly as bench.py does.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    import gpzip_amd

    c = gpzip_amd.Context()
    yield c
    c.close()
