"""Shared test set-up: import paths, the `gpu` marker, and skip logic.

Tests marked `gpu` need a gfx950 device and the built HIP library; the rest
run on the CPU (oracle pins, host logic, library load/exports).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu_lib():
    from uq_amd import lbm
    if lbm.device_count() < 1:
        pytest.fail("no gfx950 device visible to the HIP library")
    return lbm
