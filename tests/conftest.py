import os
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


def pytest_collection_modifyitems(config, items):
    # GPU tests print which case they run (a fault ends the process without a summary)
    for it in items:
        if it.get_closest_marker("gpu") is not None:
            it.user_properties.append(("gpu", True))


@pytest.fixture(scope="session")
def oracle():
    from omhd_testlib import Oracle

    return Oracle()


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu case] {request.node.nodeid}\n")
        sys.stdout.flush()
    yield
