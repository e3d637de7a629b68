"""Test configuration: the `gpu` marker, import paths, shared fixtures.

`-m "not gpu"` runs the oracle checks, host logic and the C-ABI symbol check
on the CPU; `-m gpu` runs the parity tests through the C ABI on a GPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "wgpu-voxel-engine_amd", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    import voxel_amd as V
    if V.device_count() < 1:
        pytest.fail("GPU test selected but no HIP device is visible")
    c = V.Context()
    yield c
    c.close()
