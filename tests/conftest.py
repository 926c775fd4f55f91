import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _ensure_oracle():
    import oracle
    if not oracle.LIB_PATH.exists():
        oracle.build()


_ensure_oracle()


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dd():
    """The product binding; the library must have been built."""
    import dedisp_tests_amd as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle
