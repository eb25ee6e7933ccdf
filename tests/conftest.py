import os
import sys

import pytest

# torch ships its own HIP runtime.  Importing it before the library is loaded
# makes the library bind that same runtime, so tests that hand torch device
# tensors to the library run on one runtime in the process.
try:
    import torch  # noqa: F401
except ImportError:  # CPU-only checks do not need it
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "skycube-templates_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def ctx():
    import skycube_amd

    c = skycube_amd.Context(0)
    yield c
    c.close()
