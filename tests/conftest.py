import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def demo_lines():
    """examples/demo4096.smi of the reference (4096 'SMILES NAME' lines),
    kept as a data fixture."""
    with open(os.path.join(GOLDEN, "demo4096.smi"), "rb") as f:
        return f.readlines()
