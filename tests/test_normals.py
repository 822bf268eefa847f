"""Host checks of the general-normal boundary matrices (the fix of the
reference's axis rule, kernels.cu:359-361, Issue #15): for normals off the
axes and K up to 7 (slm 120, the ASLM_MAX limit of common.cuh:7),

- the library's O(n) = <|Omega.n| Y Y'> against an independent numpy/scipy
  rule (scipy's harmonics, Gauss-Legendre split at Omega.n = 0 in a frame
  around n), 1e-12;
- the form the face kernel applies, R O(z) R^T with the packed block
  rotation R, against the same (1e-12); R orthogonal;
- the Funk-Hecke column <|Omega.n| Y_lm> (the boundary flux) against
  column 00 of the same."""
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

import transfer_amd as ta
from angular_np import even_ls, harmonics

NORMALS = [(1.0, 2.0, 3.0), (0.3, -0.1, 0.95), (-1.0, 0.5, 0.2), (0.0, 0.0, 1.0), (1.0, 1.0, 0.0), (0.0, -1.0, 0.0)]


def numpy_abs_normal(K, n, ng=None, M=None):
    n = np.asarray(n, float) / np.linalg.norm(n)
    t1 = np.cross(n, [1.0, 0, 0] if abs(n[0]) < 0.9 else [0, 1.0, 0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    ng = ng or 2 * K + 4
    M = M or 4 * K + 10
    x, w = leggauss(ng)
    mu = np.concatenate([0.5 * (x - 1), 0.5 * (x + 1)])
    wm = np.concatenate([0.5 * w, 0.5 * w])
    ph = 2 * np.pi * (np.arange(M) + 0.5) / M
    MU, PH = np.meshgrid(mu, ph, indexing="ij")
    W = (np.outer(wm, np.ones(M)) / 2 / M).ravel()
    MU, PH = MU.ravel(), PH.ravel()
    s = np.sqrt(1 - MU ** 2)
    d = MU[:, None] * n + s[:, None] * (np.cos(PH)[:, None] * t1 + np.sin(PH)[:, None] * t2)
    Y = harmonics(even_ls(K), np.clip(d[:, 2], -1, 1), np.arctan2(d[:, 1], d[:, 0]))
    return np.einsum("ip,jp,p->ij", Y * np.abs(MU), Y, W)


def unpack_rotation(K, R):
    slm = (1 + K) * (1 + 2 * K)
    D = np.zeros((slm, slm))
    s = off = 0
    for l in range(0, 2 * K + 1, 2):
        nb = 2 * l + 1
        D[s:s + nb, s:s + nb] = R[off:off + nb * nb].reshape(nb, nb)
        s += nb
        off += nb * nb
    return D


@pytest.mark.parametrize("K", [1, 3, 5, 7])
def test_normal_matrices_against_numpy(K):
    Oz = ta.abs_normal(K, (0.0, 0.0, 1.0))[0]
    for n in NORMALS:
        O, col, R = ta.abs_normal(K, n)
        want = numpy_abs_normal(K, n)
        D = unpack_rotation(K, R)
        assert np.abs(D @ D.T - np.eye(len(D))).max() < 1e-12
        np.testing.assert_allclose(O, want, rtol=0, atol=1e-12)
        np.testing.assert_allclose(D @ Oz @ D.T, want, rtol=0, atol=1e-12)
        np.testing.assert_allclose(col, want[:, 0], rtol=0, atol=1e-13)
        assert O[0, 0] == pytest.approx(0.5, abs=1e-14)      # <|Omega.n|> = 1/2
    # along the axes the general matrix is the reference's O_x, O_y, O_z (AngularData.cpp:48-66)
    a = ta.Angular(K)
    for axis, ref in enumerate((a.Ox, a.Oy, a.Oz)):
        e = np.zeros(3)
        e[axis] = 1.0
        s = a.slm
        np.testing.assert_allclose(ta.abs_normal(K, e)[0], ref.reshape(s, s).T, rtol=0, atol=1e-13)
