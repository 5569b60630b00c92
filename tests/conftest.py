"""Shared test set-up: import paths, the `gpu` marker, common fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def cb():
    import cublox_amd
    return cublox_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu_ready(cb):
    n = cb.device_count()
    if n < 1:
        pytest.fail("GPU test selected but no HIP device is visible")
    print(f"[gpu] {cb.device_name()}", flush=True)
    return True
