import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from _helpers import REF_SO, Oracle, Ref, load_pkg  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an AMD Instinct GPU (MI355X)")


@pytest.fixture(scope="session")
def orc():
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return Ref()


@pytest.fixture(scope="session")
def gb():
    return load_pkg()


@pytest.fixture(scope="session")
def ctx(gb):
    c = gb.Context(0)
    yield c
    c.close()
