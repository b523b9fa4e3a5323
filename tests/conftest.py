import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def vh():
    from __graft_entry__ import load_package
    return load_package()


@pytest.fixture(scope="session")
def ora():
    import oracle_api
    oracle_api.lib()
    return oracle_api


@pytest.fixture(scope="session")
def gpu_ready():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    torch.cuda.init()
    return True


@pytest.fixture
def opts(vh):
    """Kernel options for one test (vh_set_options), restored afterwards:
    ``opts(sir_expose=0)``."""
    saved = vh.get_options()
    yield lambda **kw: vh.set_options(**kw)
    vh.set_options(**saved)
