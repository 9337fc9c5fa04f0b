import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU test")


def _ensure_built():
    from shci_amd import _build

    if not _build.ORACLE_LIB.exists():
        _build.build_oracle()
    if not _build.HIP_LIB.exists():
        _build.build_hip()
    if not _build.HOST_LIB.exists():
        _build.build_host()


_ensure_built()


@pytest.fixture(scope="session")
def golden_dir():
    return ROOT / "tests" / "golden"
