"""Shared fixtures.  `-m gpu` tests need an MI355X and the built libssdf_amd.so; everything else
runs on the CPU (oracle, host logic, library loading)."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libssdf_amd.so)")
    config.addinivalue_line("markers", "slow: longer CPU tests")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O._load()
    return O


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; where a GPU test runs, its absence is a failure, never a skip."""
    import ssdf_amd
    L = ssdf_amd.lib()
    assert ssdf_amd.available(), "no GPU visible to libssdf_amd.so"
    return L
