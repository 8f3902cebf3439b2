import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The oracle's loops are OpenMP-parallel over every core.  Under pytest-xdist each
# worker would start one such pool, and the pools' spinning threads oversubscribe
# the cores by the worker count (a test that takes seconds alone then takes minutes),
# so each worker gets its share of the cores unless OMP_NUM_THREADS is already set.
_workers = os.environ.get("PYTEST_XDIST_WORKER_COUNT")
if _workers and "OMP_NUM_THREADS" not in os.environ:
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // max(1, int(_workers))))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def bec():
    import bec_amd
    bec_amd.lib()
    if bec_amd.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    return bec_amd
