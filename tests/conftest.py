"""Test configuration: the ``gpu`` marker and repository imports.

CPU tests (``-m "not gpu"``) check the oracle against the golden vectors,
the host logic and the C ABI's exported symbols.  GPU tests (``-m gpu``)
compare the HIP kernels, called through the C ABI, with the oracle.
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built gfx950 library")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from gm1_amd import _native

    _native.lib()  # the HIP library must load: no fallback
    return torch.device("cuda:0")
