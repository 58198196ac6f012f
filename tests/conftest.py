import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "staggered-grid_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~20 s")


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name))
    meta = json.loads(str(z["meta"]))
    return meta, {k: z[k] for k in z.files if k != "meta"}


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as O

    if not os.path.exists(O.LIB_PATH):
        import subprocess

        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return O
