import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: full-size case (seconds of CPU oracle work)")


@pytest.fixture(scope="session")
def hip():
    """The HIP library; fails (does not skip) when it cannot be loaded."""
    import hpcg_amd
    hpcg_amd.lib()
    return hpcg_amd
