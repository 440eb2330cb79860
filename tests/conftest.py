"""pytest configuration: markers and import paths.

`gpu` tests need an MI355X (HIP device) and the built libevp1d.so; all
other tests run on the CPU (oracle, golden vectors, host logic, ABI exports).
"""
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO / "cice-evp1d_amd", REPO):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _gpu_available() -> bool:
    """hipGetDeviceCount through the HIP runtime (no torch import needed)."""
    import ctypes
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
        except OSError:
            continue
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    return False


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_available():
        pytest.skip("no HIP device")
    return True
