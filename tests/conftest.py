import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: full-size parity runs")


def rel_linf(a, b):
    """max |a - b| / max |b| over one component (0/0 counts as 0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    return 0.0 if num == 0 else num / den


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sum(np.abs(b))
    num = np.sum(np.abs(a - b))
    return 0.0 if num == 0 else num / den


@pytest.fixture(scope="session")
def stage_fixtures():
    return np.load(os.path.join(GOLDEN, "stage_fixtures.npz"))


@pytest.fixture(scope="session")
def run_fixtures():
    return np.load(os.path.join(GOLDEN, "run_fixtures.npz"))


@pytest.fixture(scope="session")
def refcase_fixtures():
    return np.load(os.path.join(GOLDEN, "refcase_fixtures.npz"))


@pytest.fixture(scope="session")
def golden_index():
    import json
    with open(os.path.join(GOLDEN, "index.json")) as fh:
        return json.load(fh)
