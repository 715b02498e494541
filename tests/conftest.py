import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden_functions():
    return dict(np.load(os.path.join(GOLDEN, "functions.npz")))


@pytest.fixture(scope="session")
def golden_cfg1():
    return dict(np.load(os.path.join(GOLDEN, "cfg1.npz")))


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle_py
    return oracle_py


@pytest.fixture(scope="session")
def golden_density():
    """density.npz split by density: golden_density(2) looks like cfg1.npz."""
    d = dict(np.load(os.path.join(GOLDEN, "density.npz")))

    def part(density):
        pre = f"d{density}_"
        return {k[len(pre):]: v for k, v in d.items() if k.startswith(pre)}
    return part
