import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail, not skip, without a visible GPU: they are only selected
    with -m gpu, on a machine that must have one."""
    import torch
    assert torch.cuda.is_available(), "no GPU visible to a -m gpu test"
    from monc_amd import tvd
    tvd.load_library()
    return torch.device("cuda:0")
