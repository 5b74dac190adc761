import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_addoption(parser):
    parser.addoption("--runslow", action="store_true", default=False,
                     help="also run the long CPU oracle runs marked slow")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libwignerhip on cuda:0)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs (skipped unless --runslow)")


def pytest_collection_modifyitems(config, items):
    if config.getoption("--runslow"):
        return
    skip = pytest.mark.skip(reason="long CPU oracle run: pass --runslow")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def cuda():
    if not gpu_available():
        pytest.skip("no GPU visible")
    import torch

    return torch.device("cuda:0")
