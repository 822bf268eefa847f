"""Independent numpy/scipy evaluation of the angular integrals, used to pin
the oracle: harmonics from scipy.special.lpmv (Condon-Shortley phase removed,
the convention of Spherical.cpp:62-82) and a Gauss-Legendre x trapezoid
product rule from numpy."""
import numpy as np
from numpy.polynomial.legendre import leggauss
from scipy.special import factorial, lpmv


def harmonics(ls, mu, phi):
    """Y[(l,m)] on the given points for l in ls, m = -l..l."""
    out = []
    for l in ls:
        for m in range(-l, l + 1):
            am = abs(m)
            c = np.sqrt((2 * l + 1) * (1 if am == 0 else 2) * factorial(l - am) / factorial(l + am))
            P = lpmv(am, l, mu) * (-1.0) ** am
            out.append(c * P * (np.cos(am * phi) if m >= 0 else np.sin(am * phi)))
    return np.array(out)


def sphere_rule(n=24, M=48):
    x, w = leggauss(n)
    ph = 2 * np.pi * (np.arange(M) + 0.5) / M
    MU, PH = np.meshgrid(x, ph, indexing="ij")
    W = np.outer(w, np.ones(M)) / 2 / M
    return MU.ravel(), PH.ravel(), W.ravel()


def abs_coupling(K, axis, n=16, M=48):
    """<|Omega_axis| Y Y'> with a rule split at Omega_axis = 0 (the kink)."""
    x, w = leggauss(n)
    mu = np.concatenate([0.5 * (x - 1), 0.5 * (x + 1)])
    wm = np.concatenate([0.5 * w, 0.5 * w])
    ph = 2 * np.pi * (np.arange(M) + 0.5) / M
    MU, PH = np.meshgrid(mu, ph, indexing="ij")
    W = (np.outer(wm, np.ones(M)) / 2 / M).ravel()
    MU, PH = MU.ravel(), PH.ravel()
    s = np.sqrt(1 - MU ** 2)
    d = np.zeros((3, len(MU)))
    d[axis], d[(axis + 1) % 3], d[(axis + 2) % 3] = MU, s * np.cos(PH), s * np.sin(PH)
    Y = harmonics(even_ls(K), d[2], np.arctan2(d[1], d[0]))
    return np.einsum("ip,jp,p->ij", Y * np.abs(MU), Y, W)


def even_ls(K):
    return list(range(0, 2 * K + 1, 2))


def couplings(K, albedo=0.0, g=0.0, hg_lmax=0):
    """E[a][b][out][in] = <Y_out Omega_a W Omega_b Y_in> and
    O[n][out][in] = <|Omega_n| Y_out Y_in>, massfac[lm]."""
    mu, ph, w = sphere_rule()
    st = np.sqrt(1 - mu ** 2)
    Om = np.array([st * np.cos(ph), st * np.sin(ph), mu])
    Ye = harmonics(even_ls(K), mu, ph)
    if albedo == 0:
        E = np.einsum("op,ap,bp,ip,p->abio", Ye, Om, Om, Ye, w)
    else:
        ol = list(range(1, 2 * K + 2, 2))
        Yo = harmonics(ol, mu, ph)
        lo = np.concatenate([[l] * (2 * l + 1) for l in ol])
        gl = np.where(lo <= hg_lmax, g ** lo, 0.0)
        wo = 1 / (1 - albedo * gl)
        A = np.einsum("ep,ap,op,p->aeo", Ye, Om, Yo, w)
        E = np.einsum("aeo,o,bio->abei", A, wo, A)
    O = np.array([abs_coupling(K, n) for n in range(3)])
    le = np.concatenate([[l] * (2 * l + 1) for l in even_ls(K)])
    gle = np.where(le == 0, 1.0, np.where(le <= hg_lmax, g ** le.astype(float), 0.0))
    return E, O, 1 - albedo * gle


def reference_storage(E):
    """omega/omega_pos as AngularData.cpp:30-47 stores them, from the
    operator couplings E[a][b][out][in]: omega[a][in][out] with the first b
    whose value exceeds 1e-10 (else b = 2)."""
    slm = E.shape[-1]
    om = np.zeros((3, slm, slm))
    pos = np.zeros((3, slm, slm), np.uint32)
    for a in range(3):
        for i in range(slm):
            for o in range(slm):
                for b in range(3):
                    v = E[a, b, o, i]
                    if abs(v) > 1e-10 or b == 2:
                        om[a, i, o], pos[a, i, o] = v, b
                        break
    return om, pos
