import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    import json
    d = os.path.join(REPO, "tests", "golden")
    with open(os.path.join(d, "reference_vectors.json")) as f:
        vec = json.load(f)
    with open(os.path.join(d, "ref_numbers.json")) as f:
        nums = json.load(f)
    with open(os.path.join(d, "ref_tables.json")) as f:
        tabs = json.load(f)
    return {"vectors": vec, "numbers": nums, "tables": tabs["tables"]}
