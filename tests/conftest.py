"""Test configuration: package and oracle import paths, the ``gpu`` marker."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_PARENT = ROOT / "iron-soil-physical-simulation-algorithom-for-isaac-sim_amd"
for p in (PKG_PARENT, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def runtime():
    from isaac_bulk_hip import HipRuntime

    rt = HipRuntime("cuda:0")
    yield rt
