import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long CPU run, enabled with BQ_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("BQ_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU check; set BQ_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def gpu_lib():
    """The product library on a real device; a GPU test fails (not skips) without one."""
    import bem3_emf_qd2_amd as bq
    L = bq.lib()
    if L.bq_device_ok() != 0:
        pytest.fail("gfx950 device not usable: " + L.bq_last_error().decode())
    return bq
