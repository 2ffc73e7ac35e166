import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# The oracle's OpenMP team: the GPU host grants 16 cores to one GPU.
os.environ.setdefault("OMP_NUM_THREADS", "16")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libwatercube.so")
    config.addinivalue_line("markers", "slow: long CPU-side check")


@pytest.fixture(scope="session")
def oracle():
    from oracle import wc_oracle

    wc_oracle.lib()
    return wc_oracle


@pytest.fixture(scope="session")
def wclib():
    from watercube_amd import _lib

    _lib.lib()
    return _lib
