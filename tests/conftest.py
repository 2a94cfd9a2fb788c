import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """Skips a gpu-marked test on machines without a GPU; fails if the library
    is missing or does not load (an unresolved symbol must not read as "no
    GPU")."""
    import hairmsnn_amd as hm
    from hairmsnn_amd._lib import LIB_PATH
    if not LIB_PATH.exists():
        pytest.fail(f"HIP library missing: {LIB_PATH}")
    try:
        hm._lib.load()
    except OSError as e:
        pytest.fail(f"HIP library does not load: {e}")
    try:
        n, _ = hm.device_info()
    except Exception:
        n = 0
    if n <= 0:
        pytest.skip("no GPU visible")
    return hm


@pytest.fixture(scope="session")
def orc():
    from tests.oracle_lib import load_oracle
    return load_oracle()
