import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    with np.load(os.path.join(GOLDEN, "ref_vectors.npz")) as d:
        return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def inverse_recovery():
    rows = np.genfromtxt(os.path.join(GOLDEN, "inverse_recovery.csv"), delimiter=",",
                         names=True)
    return np.array([r["g_estimate"] for r in rows])
