import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU check")


@pytest.fixture(scope="session")
def ctx():
    from dvc_amd import _lib
    c = _lib.Context(max_iter=50)
    yield c
    c.close()
