import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(12345)


def ang_dict(a):
    """Host angular arrays of a transfer_amd.Angular as the oracle takes them."""
    return dict(omega=a.omega, omega_pos=a.omega_pos, Ox=a.Ox, Oy=a.Oy, Oz=a.Oz,
                massfac=a.massfac, slm=a.slm)


def random_vector(nP, slm, aslm, rng, smooth=False):
    f = np.zeros((nP, aslm))
    f[:, :slm] = rng.standard_normal((nP, slm))
    return f.reshape(-1)
