import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _have_gpu() -> bool:
    return Path("/dev/kfd").exists() and os.environ.get("HIP_VISIBLE_DEVICES", "x") != ""


@pytest.fixture(scope="session")
def ctx():
    """One xrh context for the whole GPU session. The library must be built (no fallback)."""
    if not _have_gpu():
        pytest.skip("no GPU visible")
    from xreg_amd import _xrh as X

    c = X.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O
