import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtk.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running case")


def has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def tk():
    return importlib.import_module("3dimagetoolkit_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


@pytest.fixture(scope="session")
def ctx(tk):
    if not has_gpu():
        pytest.skip("no GPU in this container")
    c = tk.Context(0)
    yield c
    c.close()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
