"""Shared fixtures: the CPU oracle (the checker) and the gfx950 library.

The oracle under oracle/ is loaded only from here, as the checker the HIP
results are compared against.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle
    return Oracle("f64")


@pytest.fixture(scope="session")
def oracle32():
    from oracle_lib import Oracle
    return Oracle("f32")


@pytest.fixture(scope="session")
def oracle32dp():
    """Float links with the OR, microcanonical and heatbath updates in double
    (the reference's DPUPDATES build)."""
    from oracle_lib import Oracle
    return Oracle("f32dp")


@pytest.fixture(scope="session")
def ref():
    """The reference-header harness (oracle/_ref); skipped where it is not built."""
    from oracle_lib import RefHarness
    if not RefHarness.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return RefHarness()


@pytest.fixture(scope="session")
def gf():
    import su3_ma_gauge_gpu_amd as gf
    gf._lib.load()
    return gf


@pytest.fixture(scope="session")
def gpu(gf):
    n, name = gf.device_info()
    assert n >= 1, "no HIP device"
    assert name.startswith("gfx950"), f"expected gfx950, got {name}"
    return name
