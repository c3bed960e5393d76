import os
import sys

import pytest

# torch brings its own HIP runtime; loading it before libmirage_hip makes the
# process use one runtime for both.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def handle():
    from mirage_amd import native
    h = native.Handle(0, 32768)
    yield h
    h.close()
