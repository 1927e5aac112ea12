import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gk_ref():
    import numpy as np
    z = np.load(os.path.join(GOLDEN, "gk_reference.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def is_same(a, b, tol=1e-10):
    """The reference's DataRegression<double>::is_same (tests/src/test_fixtures.h:110-122),
    vectorised: equal when both are 0, else |max-min|/|max| <= tol with max the
    operand of larger magnitude; NaN must match NaN."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nan = np.isnan(a) | np.isnan(b)
    both_nan = np.isnan(a) & np.isnan(b)
    big = np.where(np.abs(a) > np.abs(b), a, b)
    small = np.where(np.abs(a) > np.abs(b), b, a)
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.where((a == 0) & (b == 0), 0.0, (big - small) / big)
    ok = np.where(nan, both_nan, ~(rel > tol))
    return ok
