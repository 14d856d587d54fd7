import json
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcmpy_amd.so")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def ref_fixtures():
    with open(GOLDEN / "reference_fixtures.json", encoding="utf-8") as f:
        return json.load(f)


def _gpu_available() -> bool:
    try:
        from cmpy_amd import _lib

        return _lib.lib().ed_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests must run on a GPU box with the library built."""
    from cmpy_amd import _lib

    _lib.require_gpu()
    import torch

    torch.cuda.init()
    return True
