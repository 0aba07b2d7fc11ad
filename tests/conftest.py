import json
import os
import sys

import numpy as np
import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(TESTS)
sys.path.insert(0, os.path.join(ROOT, "het-scd_amd"))
sys.path.insert(0, TESTS)

GOLDEN = os.path.join(TESTS, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: large graphs (R-MAT scale 18-20)")


@pytest.fixture(scope="session")
def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def full_fixture(name):
    path = os.path.join(GOLDEN, f"full_{name}.npz")
    return np.load(path) if os.path.exists(path) else None


@pytest.fixture(scope="session")
def engine():
    """One engine context for the whole GPU session (device 0)."""
    import scd_amd
    if scd_amd.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    eng = scd_amd.Engine(device=0)
    yield eng
    eng.close()
