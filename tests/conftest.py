import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.pyoracle import REF_LIB, Reference
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref was not built (needs the SciHPC sources)")
    return Reference()


def pytest_collection_modifyitems(config, items):
    # print the test name before each GPU test so a crash names its case
    pass
