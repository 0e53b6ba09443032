import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def read_circuit(rel):
    with open(os.path.join(GOLDEN, "circuits", rel)) as f:
        return f.read()


@pytest.fixture(scope="session")
def oracle():
    from oracle import refcpu
    refcpu.lib()
    return refcpu
