import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU box)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """Initialise the library on cuda:0 once per session (with host-mapped memory)."""
    import torch
    import convfpga_amd as C
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.zeros(1, device="cuda")
    rc = C.fpga_initialize("gfx950", C.CODE_OBJECT, True)
    assert rc == 0, "fpga_initialize -> %d" % rc
    yield C
    C.fpga_final()
