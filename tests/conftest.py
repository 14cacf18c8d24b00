import json
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (run with -m gpu)")


_cache = {}


def golden(name):
    if name not in _cache:
        with open(os.path.join(GOLDEN, name + ".json")) as f:
            _cache[name] = json.load(f)
    return _cache[name]


@pytest.fixture(scope="session")
def kats():
    return golden("kats")


@pytest.fixture(scope="session")
def plans():
    return golden("plans")
