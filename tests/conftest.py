import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


def gpu_available() -> bool:
    try:
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        count = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(count)) == 0 and count.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU visible")
    from replica_exchange_wang_landau_amd.engine import default_backend

    return default_backend()


@pytest.fixture(scope="session")
def orc():
    from harness import oracle

    return oracle()


@pytest.fixture(scope="session")
def ref():
    from harness import Reference

    if not Reference.available():
        pytest.skip("compiled reference (oracle/_ref) not built here")
    return Reference()
