import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ref_tables():
    with open(os.path.join(GOLDEN, "reference_tables.json")) as f:
        return {k: np.array(v) for k, v in json.load(f).items()}


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


def rel_err(x, ref):
    """max |x - ref| / max |ref| (per array)."""
    ref = np.asarray(ref, np.float64)
    den = np.abs(ref).max()
    return float(np.abs(np.asarray(x, np.float64) - ref).max() / (den if den > 0 else 1.0))
