import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(params=[0, 1], ids=["sum_reference", "sum_tree"])
def sum_order(request):
    """Summation order of the solver reductions, set in the oracle (og_set_sum_mode) for the
    test's duration: 0 = the reference's sequential sums, 1 = double-double (the GPU library's
    FUGU_SUM_TREE).  The value is also the GPU library's order constant."""
    import pyoracle as O
    prev = O.get_sum_mode()
    O.set_sum_mode(request.param)
    yield request.param
    O.set_sum_mode(prev)
