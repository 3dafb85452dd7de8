import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "npb-gpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_cases.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN) as f:
        doc = json.load(f)
    return {c["name"]: c for c in doc["cases"]}


@pytest.fixture(scope="session")
def npbcg_lib():
    import npbcg  # raises ImportError if the HIP library is not built

    return npbcg


@pytest.fixture(scope="session")
def gpu(npbcg_lib):
    """The HIP library with at least one visible GPU; fails (not skips) otherwise."""
    n = npbcg_lib.device_count()
    if n < 1:
        pytest.fail("no GPU visible to libnpbcg: GPU tests must run on a GPU host")
    return npbcg_lib
