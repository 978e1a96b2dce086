import os
import sys

import numpy as np
import pytest

# One HIP runtime per process: torch bundles its own libamdhip64, and it must
# be loaded before the engine binds one (see tapkee_amd/_lib.py).
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def engine():
    import tapkee_amd
    tapkee_amd._lib.lib()
    return tapkee_amd


def rng(seed=0):
    return np.random.Generator(np.random.PCG64(seed))
