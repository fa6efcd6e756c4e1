import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libmcl.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One libmcl context for the whole GPU session (the HIP path, never a fallback)."""
    from mcluminescence_amd._abi import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
