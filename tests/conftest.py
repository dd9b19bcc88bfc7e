import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PARAMS = os.path.join(ROOT, "phylocsf_amd", "data", "PhyloCSF_Parameters")
EXAMPLES = os.path.join(ROOT, "tests", "golden", "PhyloCSF_Examples")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def params_dir():
    return PARAMS


@pytest.fixture(scope="session")
def examples_dir():
    return EXAMPLES
