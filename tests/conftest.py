import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.jet_oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle.jet_oracle import Ref, have_ref

    if not have_ref():
        pytest.skip("oracle/_ref/libjetref.so not built (needs the reference tree)")
    return Ref()
