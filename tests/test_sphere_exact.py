"""The exact sphere solution (tests/sphere_exact.py) against closed forms:
the centre of the two-region sphere, a homogeneous sphere's centre, the
optically thin far field (int chord dOmega = V / r^2 for r >> a), the
infinite-medium limit, and a brute-force angular integral."""
import numpy as np

import sphere_exact as se


def test_centre_closed_form():
    # from r = 0 every ray crosses the core (a) then the shell (R - a)
    a, R, kc, ks, ic, is_ = 0.3, 1.0, 11.0, 1.0, 1.0, 0.5
    want = 4 * np.pi * (ic * (1 - np.exp(-kc * a)) + is_ * (1 - np.exp(-ks * (R - a))) * np.exp(-kc * a))
    got = se.U(0.0, a=a, R=R, kc=kc, ic=ic, ks=ks, is_=is_)
    assert abs(got - want) <= 1e-12 * want


def test_infinite_medium_limit():
    # a thick homogeneous sphere: U -> 4 pi I_p deep inside
    got = se.U(0.1, a=0.5, R=1.0, kc=200.0, ic=1.0, ks=200.0, is_=1.0)
    assert abs(got - 4 * np.pi) <= 1e-9 * 4 * np.pi


def test_thin_far_field():
    # thin core, vacuum shell: U(r) -> kappa I V / r^2 far away
    a, kc, r = 0.01, 1e-3, 0.9
    got = se.U(r, a=a, R=1.0, kc=kc, ic=1.0, ks=0.0, is_=0.0, npts=400)
    want = kc * (4 / 3 * np.pi * a ** 3) / r ** 2
    assert abs(got - want) <= 1e-3 * want


def test_brute_force_directions():
    # a plain midpoint rule in theta of the same integrand (it converges
    # slowly at the core's tangent directions: 1.2e-7 relative here)
    r, N = 0.45, 200000
    th = (np.arange(N) + 0.5) * np.pi / N
    I = se.intensity(r, np.cos(th), 0.3, 1.0, 11.0, 1.0, 1.0, 0.0)
    brute = 2 * np.pi * np.sum(I * np.sin(th)) * np.pi / N
    assert abs(se.U(r, npts=800) - brute) <= 1e-6 * brute
