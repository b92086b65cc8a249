"""Shared test setup: import paths, the `gpu` marker and common fixtures."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for sub in ("sophus-rs_amd", "oracle"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def sg():
    import sophus_opt_gpu

    sophus_opt_gpu.library()
    return sophus_opt_gpu


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
