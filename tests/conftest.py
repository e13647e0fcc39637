import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP engine)")


@pytest.fixture(scope="session")
def ref_data():
    """The reference's own test datasets (data/*.csv parsed with strtof)."""
    z = np.load(os.path.join(GOLDEN, "reference_data.npz"))

    def get(name):
        return z[name + "__coords"], z[name + "__weights"]

    return get


@pytest.fixture(scope="session")
def ref_truth():
    return np.load(os.path.join(GOLDEN, "reference_truth.npz"))


@pytest.fixture(scope="session")
def config1_golden():
    return np.load(os.path.join(GOLDEN, "config1_golden.npz"))
