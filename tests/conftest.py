import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "geodynamo.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


try:  # torch's bundled HIP runtime must be loaded before libgdhip's
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O
