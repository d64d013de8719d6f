import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU test, skipped unless GW_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("GW_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow; set GW_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(autouse=True)
def _threads():
    import torch

    torch.set_num_threads(min(8, os.cpu_count() or 1))
