"""Test configuration: the `gpu` marker, import paths and shared helpers."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
# NE_ORACLE_DIR: a directory holding a modified copy of oracle.py, which the mutation test
# (test_oracle_mutants.py) runs the pins against in a child process.
ORACLE_DIR = Path(os.environ.get("NE_ORACLE_DIR", REPO / "oracle"))
for p in (REPO / "numericalearth.jl_amd", ORACLE_DIR, REPO / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device is visible")
