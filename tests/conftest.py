import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfwx on cuda:0)")
    config.addinivalue_line("markers", "slow: large sizes (minutes of CPU oracle time)")


@pytest.fixture(scope="session")
def kats():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "kats.json")) as f:
        return json.load(f)["kats"]


@pytest.fixture(scope="session")
def gen_ref():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "gen_ref.json")) as f:
        return json.load(f)["cases"]
