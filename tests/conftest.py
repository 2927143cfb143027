import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "go-melt_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


# PyTorch ships its own HIP runtime under the same soname as the system one
# libgomelt.so links, and the first one loaded serves the whole process;
# PyTorch does not start on the system one.  The tests of the learned
# operator share the process with the other GPU tests, so PyTorch is loaded
# first (a CPU-only PyTorch loads no HIP runtime).
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libgomelt.so")
    config.addinivalue_line("markers", "slow: long-running (skipped unless selected)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import gomelt_oracle as O
    O.set_precision("float64")
    return O
