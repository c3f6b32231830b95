import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """The HIP path must run: a GPU test fails (never skips) when no device is usable."""
    import torch

    from mpmc_hnu_amd import native

    native.library()  # raises when the gfx950 library is not built
    assert torch.cuda.is_available(), "gpu test requires a visible GPU"
    return torch.device("cuda", 0)
