import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def O64():
    from oracle import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def O32():
    from oracle import Oracle
    return Oracle(32)


GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def prog():
    """Program vectors of the reference on the 48-pixel synthetic set."""
    return np.load(os.path.join(GOLDEN, "programs.npz"))


@pytest.fixture(scope="session")
def ctx():
    from cryoim_amd import lib
    # torch carries its own copy of the HIP runtime: it must reach the GPU
    # before the library's runtime does, or it finds none
    try:
        import torch
        torch.cuda.init()
    except Exception:
        pass
    c = lib.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
