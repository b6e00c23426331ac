import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build(ref=False)
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a visible GPU (tests marked gpu only)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import cudagsea_amd
    cudagsea_amd.lib()
    return cudagsea_amd
