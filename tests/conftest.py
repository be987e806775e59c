import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for every GPU test (created once, in this process)."""
    from wadjet_amd.gpu import Context
    c = Context(0)
    print(f"[ctx] device 0: {c.cus} CUs, {c.hbm >> 30} GiB, LDS/block {c.lds}", flush=True)
    yield c
    c.close()
