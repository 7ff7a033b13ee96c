"""pytest configuration: the ``gpu`` marker and shared helpers."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: a test marked gpu fails loudly if the library or device is missing."""
    import torch

    from somax_amd import _lib

    assert torch.cuda.is_available(), "gpu test run without a GPU"
    _lib.load()
    _lib.init_device()
    return torch.device("cuda:0")
