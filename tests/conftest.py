import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "aocl-data-analytics_amd"))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X)")


@pytest.fixture(scope="session")
def known():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def nn_csv():
    out = {}
    for name in ("2092x5", "2125x20"):
        z = np.load(os.path.join(GOLDEN, f"nn_csv_{name}.npz"), allow_pickle=False)
        out[name] = {k: z[k] for k in z.files}
    return out


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    return o
