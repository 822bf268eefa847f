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


_REF_FIX = os.path.join(ROOT, "tests", "golden", "ref_angular.npz")


def ref_angular_fixture(K):
    """The reference's own AngularData arrays for K (1..7) from the committed
    fixture (tests/golden/make_ref_fixtures.py wrote it from the reference's
    sources), as the oracle takes them."""
    z = np.load(_REF_FIX)
    s = (1 + K) * (1 + 2 * K)
    return dict(omega=z["K%d_omega" % K], omega_pos=z["K%d_omega_pos" % K], Ox=z["K%d_Ox" % K],
                Oy=z["K%d_Oy" % K], Oz=z["K%d_Oz" % K], slm=s)


def ref_xE1_fixture():
    z = np.load(_REF_FIX)
    return z["xE1_x"], z["xE1"]
