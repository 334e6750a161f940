import json
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden(name: str) -> list[dict]:
    return json.loads((GOLDEN / name).read_text())["cases"]


@pytest.fixture(scope="session")
def ctx():
    import synchrowords_amd as sw
    c = sw.Context(0)
    yield c
    c.close()
