import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def pytest_collection_modifyitems(config, items):
    # GPU tests need a device; on a CPU-only host they are skipped. A missing or
    # broken library is not a reason to skip: those tests then fail loudly.
    try:
        from soil_amd import device_count
        have = device_count() > 0
    except Exception:
        return
    if have:
        return
    skip = pytest.mark.skip(reason="no HIP device on this host")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
