import json
import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden_paillier.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def H(s: str) -> int:
    return int(s, 16)


@pytest.fixture(scope="session")
def ctx():
    """The process-wide device context (GPU tests only)."""
    from cuhepy_amd.engine import Context
    return Context.default(0)
