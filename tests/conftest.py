import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libemsa.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def pq_pairs():
    """The reference's data/P*.txt, data/Qnew*.txt pairs (row correspondence,
    one exact rigid transform, Qnew = A P + b)."""
    from cuda_emicp_softassign_amd.clouds import load_xyz
    out = {}
    for n in (101, 1007):
        out[n] = (load_xyz(os.path.join(GOLDEN, "P%d.txt" % n)),
                  load_xyz(os.path.join(GOLDEN, "Qnew%d.txt" % n)))
    return out
