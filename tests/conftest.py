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
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libdwfhip.so")


@pytest.fixture
def rng():
    return np.random.default_rng(20240611)


def print_case(**kw):
    """Name the case (flushed) before a GPU step, so a failing run says which one."""
    print("CASE " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)
