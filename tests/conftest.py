import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C-ABI")


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np

    gd = os.path.join(HERE, "golden")
    with open(os.path.join(gd, "fdmt_golden.json")) as f:
        man = json.load(f)
    arr = np.load(os.path.join(gd, "fdmt_golden.npz"), allow_pickle=False)
    return man, arr
