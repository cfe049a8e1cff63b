import os
import sys

import pytest

# torch ships its own HIP runtime (libamdhip64.so.7); loading it before
# librajzel_amd.so makes every test in the session share that one runtime.
try:
    import torch  # noqa: F401
except Exception:  # torch is optional for everything but the autograd tests
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the C ABI")


@pytest.fixture(scope="session")
def ora():
    from oracle.bindings import Restatement
    return Restatement()


@pytest.fixture(scope="session")
def ora64():
    from oracle.bindings import Restatement
    return Restatement(f64=True)


@pytest.fixture(scope="session")
def ref():
    from oracle.bindings import Reference
    if not Reference.available():
        pytest.skip("reference build oracle/_ref not present")
    return Reference()
