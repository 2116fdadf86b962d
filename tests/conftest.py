import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
EXPLICIT = os.path.join(GOLDEN, "explicit_data")
TEST_MODELS = os.path.join(GOLDEN, "test_models")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def golden(name):
    return np.load(os.path.join(EXPLICIT, name), allow_pickle=False)


@pytest.fixture(scope="session")
def explicit_dir():
    return EXPLICIT


@pytest.fixture(scope="session")
def test_models_dir():
    return TEST_MODELS
