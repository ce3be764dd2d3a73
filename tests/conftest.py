import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if HERE not in sys.path:
    sys.path.insert(0, HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the device)")
    config.addinivalue_line("markers", "slow: longer CPU check")


@pytest.fixture(scope="session")
def toy():
    with open(os.path.join(HERE, "golden", "toy_example.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def crosschecks():
    with open(os.path.join(HERE, "golden", "crosscheck_moves.json")) as f:
        return json.load(f)
