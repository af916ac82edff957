import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuda-plasticity_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")
    config.addinivalue_line("markers", "full_size: full-length runs against the C oracle (minutes of host CPU); "
                            "run with CDD_RUN_FULL=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("CDD_RUN_FULL") == "1":
        return
    skip = pytest.mark.skip(reason="full-size oracle run: set CDD_RUN_FULL=1")
    for item in items:
        if "full_size" in item.keywords:
            item.add_marker(skip)
