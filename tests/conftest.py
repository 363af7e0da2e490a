import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle.load_oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import oracle
    r = oracle.load_reference()
    if r is None:
        pytest.skip("reference C not built (oracle/_ref/libref.so absent)")
    return r


@pytest.fixture(scope="session")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")
