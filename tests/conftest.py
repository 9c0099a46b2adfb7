import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libqrhip.so")
    config.addinivalue_line("markers", "slow: full-size workload checks")


@pytest.fixture(scope="session")
def hamiltonians():
    with open(os.path.join(GOLDEN, "hamiltonians.json")) as f:
        ham = json.load(f)
    return {k: (v["labels"], [complex(a, b) for a, b in v["coeffs"]]) for k, v in ham.items()}


@pytest.fixture(scope="session")
def csr_digests():
    with open(os.path.join(GOLDEN, "csr_digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def config1():
    with open(os.path.join(GOLDEN, "config1.json")) as f:
        cfg = json.load(f)
    cfg["coeffs"] = [complex(a, b) for a, b in cfg["coeffs"]]
    cfg["y"] = np.load(os.path.join(GOLDEN, "config1_y.npy"))
    return cfg
