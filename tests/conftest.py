import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def golden_functions():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "functions.npz"))


@pytest.fixture(scope="session")
def golden_search():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "search.npz"))
