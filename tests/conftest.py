import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np
    d = os.path.join(ROOT, "tests", "golden")
    runs = np.load(os.path.join(d, "reference_runs.npz"))
    meta = json.load(open(os.path.join(d, "reference_meta.json")))
    files = json.load(open(os.path.join(d, "reference_files.json")))
    return {"runs": runs, "meta": meta, "files": files}
