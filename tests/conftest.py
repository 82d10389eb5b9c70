import os
import sys

import numpy as np
import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
PKG = os.path.join(REPO, "16s-central-folding_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (PKG, os.path.join(REPO, "oracle"), GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)

import tis_ctypes as T  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def lib():
    return T.load_library()


@pytest.fixture(scope="session")
def cfg1(lib):
    """16SCD, 5 mM MgCl2 + 50 mM KCl, 350 A, T = 0.596, seed 1 (3,659 particles)."""
    d, draws = T.build_16scd(lib, T=0.596, c_mg=0.005, c_k=0.050, seed=1)
    return d, draws


@pytest.fixture(scope="session")
def golden_cfg1():
    return dict(np.load(os.path.join(GOLDEN, "ref_cfg1.npz")))


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
