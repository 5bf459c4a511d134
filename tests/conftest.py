import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def oracle():
    import refcheck
    return refcheck.Oracle()


@pytest.fixture(scope="session")
def ref():
    import refcheck
    if not refcheck.ref_available():
        pytest.skip("oracle/_ref (the reference's CPU backend built from its sources) is not built")
    return refcheck.Ref()
