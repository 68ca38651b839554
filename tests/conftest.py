import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def pytest_collection_modifyitems(session, config, items):
    """The C++ programs of test_gpu_cpp_interface.py run as child processes and launch
    cooperative kernels; in our runs such launches failed with hipErrorIllegalState once the
    test process itself had launched cooperative kernels. Those tests run first."""
    first = [it for it in items
             if os.path.basename(str(it.fspath)) == "test_gpu_cpp_interface.py" and it.name.startswith("test_cpp_")]
    if first:
        rest = [it for it in items if it not in first]
        items[:] = first + rest
