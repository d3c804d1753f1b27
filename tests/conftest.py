import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libebi_hip.so")


@pytest.fixture(scope="session")
def ctx():
    import ebi_amd

    c = ebi_amd.Context(0)
    yield c
    c.close()
