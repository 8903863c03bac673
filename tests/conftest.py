import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    """Relative L2 error ||a-b|| / ||b|| (the reference's d3error measure)."""
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    if not os.path.exists(oracle.SO):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    return oracle


@pytest.fixture(scope="session")
def gpu_lib():
    import elfmmlib3d_amd as e
    e.library()        # raises if the HIP library is missing
    return e
