import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "matter-tests_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import mpm_abi  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long physics runs")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (test infrastructure). Built by __graft_entry__.build()."""
    return mpm_abi.oracle_library()


@pytest.fixture(scope="session")
def engine():
    """The HIP engine. A GPU test fails loudly when it is not built or no device is present."""
    lib = mpm_abi.product_library()
    return lib


def rel_err(a, b, floor=0.0):
    """Norm-wise relative error max|a-b| / max(max|b|, floor) over the finite entries of one field.
    NaN entries must sit at the same places in both (the reference produces NaN in some states, e.g. the
    mu(I)-MCC viscosity under tension); otherwise the error is infinite. `floor` is the field's physical
    scale, so that a field that is pure rounding noise (B of a uniform flow, strains of an undeformed
    body) is measured against its natural size instead of against noise."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.size == 0:
        return 0.0
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return float("inf")
    m = ~nb
    if not m.any():
        return 0.0
    scale = max(float(np.max(np.abs(b[m]))), floor, 1e-300)
    return float(np.max(np.abs(a[m] - b[m]))) / scale
