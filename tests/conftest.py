import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "oidn-protoray_amd", os.path.join("oidn-protoray_amd", "scenes")):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def results_dir():
    """Directory for the JSON / npz records GPU tests leave behind (PRT_RESULTS_DIR), or None:
    then nothing is written."""
    d = os.environ.get("PRT_RESULTS_DIR")
    if d:
        os.makedirs(d, exist_ok=True)
    return d


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
