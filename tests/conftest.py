import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "ref: needs the reference checker oracle/_ref/libgzref.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle.bindings import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.bindings import Reference, have_reference
    if not have_reference():
        pytest.skip("oracle/_ref/libgzref.so not built (reference sources absent at build time)")
    return Reference()
