import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built HIP library; calls through the C ABI")


@pytest.fixture(scope="session")
def gpu():
    """The device for gpu-marked tests: fails loudly (never skips) when the HIP path cannot run."""
    import torch

    from peregrine_amd import native

    if not torch.cuda.is_available():
        pytest.fail("a gpu test ran without a visible GPU")
    native.lib()
    return torch.device("cuda", 0)
