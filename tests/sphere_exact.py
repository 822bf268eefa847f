"""The exact transport solution of a purely absorbing two-region sphere (the
checker of tests/test_gpu_sphere.py; test infrastructure).

Core r < a with (kappa_c, I_c), shell a < r < R with (kappa_s, I_s), vacuum
outside.  Along the backward ray y(s) = x - s Omega from a point at radius r
(mu = cos of Omega against the radial direction), |y|^2 = r^2 - 2 r mu s +
s^2, so the ray meets a sphere of radius rho at s = r mu +- sqrt(rho^2 -
r^2 (1 - mu^2)).  With piecewise constant data the intensity is the sum over
the ray's segments of I_seg (1 - exp(-kappa_seg len_seg)) exp(-tau before
the segment), and

    U(r) = 2 pi int_{-1}^{1} I(r, mu) dmu,

integrated by Gauss-Legendre on the pieces between the directions tangent
to the core (where the integrand has square-root behaviour).
"""
import numpy as np


def intensity(r, mu, a, R, kc, ic, ks, is_):
    """I(r, mu) for arrays mu (backward rays from radius r <= R)."""
    mu = np.asarray(mu, float)
    b2 = r * r * (1.0 - mu * mu)
    s_exit = r * mu + np.sqrt(np.maximum(R * R - b2, 0.0))
    hit = a * a - b2 > 0
    sq = np.sqrt(np.maximum(a * a - b2, 0.0))
    s1 = np.where(hit, np.clip(r * mu - sq, 0.0, s_exit), s_exit)
    s2 = np.where(hit, np.clip(r * mu + sq, 0.0, s_exit), s_exit)
    # segments: shell [0, s1), core [s1, s2), shell [s2, s_exit)
    l0, l1, l2 = s1, s2 - s1, s_exit - s2
    t0 = ks * l0
    t1 = t0 + kc * l1
    out = is_ * (1.0 - np.exp(-ks * l0))
    out = out + ic * (1.0 - np.exp(-kc * l1)) * np.exp(-t0)
    out = out + is_ * (1.0 - np.exp(-ks * l2)) * np.exp(-t1)
    return out


def U(r, a=0.3, R=1.0, kc=11.0, ic=1.0, ks=1.0, is_=0.0, npts=200):
    """U(r) = int I dOmega at radius r (0 <= r <= R)."""
    cuts = [-1.0, 1.0]
    if r > a:
        mt = np.sqrt(1.0 - (a / r) ** 2)
        cuts = [-1.0, mt, 1.0]
    x, w = np.polynomial.legendre.leggauss(npts)
    tot = 0.0
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        m = 0.5 * (hi - lo) * x + 0.5 * (hi + lo)
        tot += 0.5 * (hi - lo) * np.dot(w, intensity(r, m, a, R, kc, ic, ks, is_))
    return 2.0 * np.pi * tot


def U_table(rs, **kw):
    return np.array([U(float(r), **kw) for r in rs])
