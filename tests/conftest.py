import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden(name):
    import numpy as np
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    d = {k: z[k] for k in z.files}
    cfg = {k[4:]: (d[k].item() if d[k].shape == () else d[k]) for k in d if k.startswith("cfg_")}
    d["cfg"] = cfg
    return d


@pytest.fixture(scope="session")
def golden():
    return load_golden
