import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REF_ROOT = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libcs3d.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gp30():
    from cellsim3d_amd import configs
    return configs.unit_mesh()


@pytest.fixture(scope="session")
def oracle_mesh(gp30):
    from cellsim3d_amd import configs, mesh as M
    from oracle import oracle as O
    return O.build_mesh(M.tables(gp30), configs.rest_positions(gp30))


def ref_available():
    return os.path.exists(os.path.join(REF_ROOT, "C180_original"))


def relerr(a, b):
    """max |a-b| / max(|b|) over a field (scale-relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-30)
    return float(np.max(np.abs(a - b)) / den)


def elemerr(a, b, floor_frac=1e-6):
    """max over elements of |a-b| / max(|b|, floor_frac * max|b|): the relative
    error of each element, with values far below the field's scale measured
    against that floor instead of themselves."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    scale = max(float(np.max(np.abs(b))), 1e-30)
    den = np.maximum(np.abs(b), floor_frac * scale)
    return float(np.max(np.abs(a - b) / den))
