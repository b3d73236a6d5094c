import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN_DIR, "reference_goldens.json")) as f:
        return json.load(f)


def load_acor_fixture(n: int):
    """The reference's acor fixtures as 2 x n matrices (amd_gcn_test.clj:322-342):
    row 0 the data, row 1 the data x2 (67, 112640) or +1 (367), computed in
    double and rounded to float as neanderthal's transfer! does."""
    import numpy as np
    with open(os.path.join(GOLDEN_DIR, f"acor-data-{n}")) as f:
        d = np.array([np.float32(float(t)) for t in f.read().split()], np.float32)
    d64 = d.astype(np.float64)
    row1 = (d64 + 1.0) if n == 367 else (d64 * 2.0)
    return np.stack([d, row1.astype(np.float32)], 1)
