import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "thesis-source-code_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

# Under pytest-xdist every worker loads the oracle's OpenMP library, whose
# teams would each start one thread per core; their busy-waiting threads then
# slow one another down by orders of magnitude.  Split the cores between the
# workers unless the caller chose a thread count.
if os.environ.get("PYTEST_XDIST_WORKER") and "OMP_NUM_THREADS" not in os.environ:
    _workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1") or 1)
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // max(1, _workers)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def vti():
    import vtiamd

    return vtiamd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx(vti):
    c = vti.Context(0)
    yield c
    c.close()
