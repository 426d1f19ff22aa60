from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:  # shared helpers (oracle_check)
    sys.path.insert(0, str(TESTS))

# The library replays images below 2^18 pixels with L2 gathers by default;
# the test fields are small, so the session lets the LDS-box and hybrid
# replays run at any size (tests of the defaults set these back to 1).
os.environ.setdefault("RLIC_REPLAY_LDS", "2")


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X / gfx950)")


def _gpu_available() -> bool:
    try:
        import torch
    except ImportError:
        return False
    return bool(torch.cuda.is_available())


@pytest.fixture(scope="session")
def gpu_device():
    """The ROCm device the GPU tests run on; fails (not skips) without one."""
    import torch

    if not _gpu_available():
        pytest.fail("GPU test selected but no ROCm device is visible")
    return torch.device("cuda:0")
