import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "soverign-gpu-accelerator-_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblakshya.so on the device)")
    config.addinivalue_line("markers", "slow: full-size workloads")


@pytest.fixture(scope="session")
def native():
    from lakshya import native as N
    return N


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def gpu(native):
    if native.device_count() < 1:
        pytest.fail("GPU test without a visible device")
    return native


def lp_from_rows(rows, b, c, ncols):
    """Small helper: LP from dense row lists."""
    import scipy.sparse as sp
    A = sp.csc_matrix(np.array(rows, dtype=float))
    A.sort_indices()
    from lakshya.gen import LP
    return LP(A.shape[0], ncols, A.indptr.astype(np.int32), A.indices.astype(np.int32),
              A.data.astype(np.float64), np.asarray(b, float), np.asarray(c, float), "adhoc")
