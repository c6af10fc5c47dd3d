import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "quarry-gpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def qabi():
    """The product's C-ABI binding; fails loudly when the HIP library is missing."""
    from quarry import _abi

    _abi.lib()
    return _abi


@pytest.fixture(scope="session")
def scenes():
    from quarry import scenes

    return scenes
