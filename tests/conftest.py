import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ramanujandreams-cuda_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP walk engine)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def published():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "published_records.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def engine():
    """The HIP walk engine through its C ABI (GPU tests only)."""
    import cmfwalk

    cmfwalk.load()
    return cmfwalk
