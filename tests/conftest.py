import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs (minutes each); run with GTO_RUN_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("GTO_RUN_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="long CPU-oracle run; set GTO_RUN_SLOW=1")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def built():
    """The native library and the oracle must be built (by __graft_entry__.build())."""
    import gtopen_amd
    import oracle

    gtopen_amd.lib()
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    oracle.lib()
    return True


def golden(name):
    return os.path.join(ROOT, "tests", "golden", name)
