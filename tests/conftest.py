import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def _gpu_available() -> bool:
    try:
        import hairmsnn_amd as hm
        n, _ = hm.device_info()
        return n > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips a gpu-marked test on machines without a GPU; fails if the library is missing."""
    import hairmsnn_amd as hm
    from hairmsnn_amd._lib import LIB_PATH
    if not LIB_PATH.exists():
        pytest.fail(f"HIP library missing: {LIB_PATH}")
    if not _gpu_available():
        pytest.skip("no GPU visible")
    return hm


@pytest.fixture(scope="session")
def orc():
    from tests.oracle_lib import load_oracle
    return load_oracle()
