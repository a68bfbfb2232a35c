"""Test configuration: the ``gpu`` marker and shared helpers.

``-m "not gpu"`` runs the oracle pinning, host planning and C-ABI export checks on CPU;
``-m gpu`` runs the HIP kernels against the oracle on a ROCm device.
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu_device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible ROCm device")
    return torch.device("cuda", 0)
