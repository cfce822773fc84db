import os
import sys

import pytest

# Under pytest-xdist every worker would start one OpenMP thread per CPU for the
# oracle's loops, and the spinning threads of several workers oversubscribe the
# machine many times over (minutes per test instead of a second): give each
# worker its share of the CPUs, before anything loads the OpenMP runtime.
if "PYTEST_XDIST_WORKER_COUNT" in os.environ:
    _workers = int(os.environ["PYTEST_XDIST_WORKER_COUNT"])
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // _workers))

ROOT =os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from __graft_entry__ import load_package  # noqa: E402

load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")
