"""Shared test setup: import paths, the `gpu` marker and common fixtures."""
import os
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


def _host_threads() -> int:
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


@pytest.fixture(scope="session")
def orc():
    """The oracle on this host's cores: its threaded loops give bitwise the result of its
    sequential ones (test_threaded_oracle_is_bitwise_sequential)."""
    import oracle

    for variant in oracle.VARIANTS:
        oracle.set_threads(min(16, _host_threads()), variant=variant)
    return oracle
