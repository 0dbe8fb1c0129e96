import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session", autouse=True)
def _one_blas_thread():
    """The CPU tests factor small dense matrices (scipy's Radau, numpy
    solves): one BLAS thread each, so that threads spinning beside a busy
    host (a parallel build, pytest-xdist workers) do not slow them down a
    hundredfold."""
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        yield


@pytest.fixture(scope="session")
def cb6r5():
    from cmaq_amd.ros3 import load_mechanism
    return load_mechanism("cb6r5_ae7_aq")


@pytest.fixture(scope="session")
def cracmm2():
    from cmaq_amd.ros3 import load_mechanism
    return load_mechanism("cracmm2")


@pytest.fixture(scope="session")
def oracle_cb6r5(cb6r5):
    from oracle.oracle import Oracle
    return Oracle(cb6r5)


@pytest.fixture(scope="session")
def oracle_cracmm2(cracmm2):
    from oracle.oracle import Oracle
    return Oracle(cracmm2)


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
