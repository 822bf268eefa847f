"""Pins the oracle's post-processing (MeshData.cpp:96-136) and the boundary
flux by checks independent of its formulas:

- the moments T_ab = 4 pi <Omega_a Omega_b psi> of psi = sum f_lm Y_lm,
  recomputed by a Gauss-Legendre x trapezoid sphere rule over scipy's
  harmonics, and their trace U;
- the flux W of ComputeFlux on a T field linear in space, which P1
  differentiates exactly: with s_j = 3 V grad(e_j) (this library's
  generator convention) the reference's formula gives W = +(1/kappa) div T,
  i.e. minus the physical flux -(1/kappa) div T;
- the outward boundary flux summed over the boundary against the lm = 0
  rows of the operator (the stiffness rows sum to zero for the constant
  test function, so what remains is mass and boundary);
- the Milne slab: the solved emergent flux at x = +-L and the solved W
  against the exact slab flux 2 pi (E3(d1) - E3(d2)) (main.cpp:505-508 is
  the half-space form 2 pi E3)."""
import numpy as np
import pytest
from scipy.special import expn

import oracle
import transfer_amd as ta
from angular_np import even_ls, harmonics, sphere_rule
from conftest import ang_dict
from reflect import masked_cg, reflect_mask


def _problem(K, albedo=0.0, n=3):
    m = ta.Mesh.box((n, n + 1, n + 2), (0, 0, 0), (1.0, 1.3, 1.7), jitter=0.2, seed=5)
    m.set_fields(kappa=m.field("uniform", [0.5, 2.0], seed=3), Ip=m.field("const", [1.0]))
    return m, ta.Angular(K, albedo)


@pytest.mark.parametrize("K", [1, 2, 3])
def test_moments_against_sphere_quadrature(K):
    m, a = _problem(K)
    op = oracle.OracleProblem(m.arrays(), ang_dict(a))
    rng = np.random.default_rng(K)
    f = np.zeros((op.nP, op.aslm))
    f[:, :op.slm] = rng.standard_normal((op.nP, op.slm))
    U, T = op.moments(f.ravel())
    mu, ph, w = sphere_rule(n=2 * K + 6, M=4 * K + 12)
    st = np.sqrt(1 - mu ** 2)
    Om = np.array([st * np.cos(ph), st * np.sin(ph), mu])
    Y = harmonics(even_ls(K), mu, ph)                 # [slm][points]
    psi = f[:, :op.slm] @ Y                          # [nP][points]
    for a in range(3):
        for b in range(3):
            want = 4 * np.pi * psi @ (Om[a] * Om[b] * w)
            np.testing.assert_allclose(T[3 * a + b], want, rtol=0, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(T[0] + T[4] + T[8], U, atol=1e-12 * np.abs(U).max())
    np.testing.assert_allclose(U, 4 * np.pi * psi @ w, atol=1e-12 * np.abs(U).max())


def test_flux_of_a_linear_moment_field():
    m, a = _problem(1)
    ar = m.arrays()
    op = oracle.OracleProblem(ar, ang_dict(a))
    rng = np.random.default_rng(9)
    A = rng.standard_normal((3, 3))
    B = rng.standard_normal((3, 3, 3))              # T_aj(x) = A_aj + B_ajc x_c, not symmetric in (a, j)
    x = ar["coords"]
    T = (A[:, :, None] + np.einsum("ajc,vc->ajv", B, x)).reshape(9, op.nP)
    W = op.flux(T)
    kappa = np.array([t["kappa_volume"] for t in ar["mesh"]])
    tets = np.array([t["p"] for t in ar["mesh"]])
    P = x[tets]
    vol = np.abs(np.einsum("ti,ti->t", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])) / 6
    kappa = kappa / vol
    divT = np.einsum("aja->j", B)                    # (div T)_j = d_a T_aj
    np.testing.assert_allclose(W, divT[:, None] / kappa[None, :], rtol=1e-11)


@pytest.mark.parametrize("K,albedo", [(1, 0.0), (3, 0.5)])
def test_boundary_flux_closes_the_lm0_balance(K, albedo):
    m, a = _problem(K, albedo)
    m.set_face_kind_plane(0, 0.0, 1)                 # one reflecting plane carries no flux
    ar = m.arrays()
    op = oracle.OracleProblem(ar, ang_dict(a))
    rng = np.random.default_rng(2)
    f = np.zeros((op.nP, op.aslm))
    f[:, :op.slm] = rng.standard_normal((op.nP, op.slm))
    F = op.boundary_flux(f.ravel())
    bnd, kind = ar["bnd"], ar["face_kind"]
    surf = np.abs(bnd["s"].sum(axis=1))
    assert np.all(F[kind == 1] == 0) and (kind == 1).any()
    Af = op.lhs(f.ravel()).reshape(op.nP, op.aslm)
    tets = np.array([t["p"] for t in ar["mesh"]])
    kv = np.array([t["kappa_volume"] for t in ar["mesh"]])
    mass00 = np.sum(kv * (1 - albedo) / 4 * f[tets, 0].sum(axis=1))
    want = 4 * np.pi * (Af[:, 0].sum() - mass00)
    assert abs(np.sum(surf * F) - want) <= 1e-11 * np.sum(np.abs(surf * F))


def _slab_flux(x, L):
    """Exact net flux of the slab [-L, L], kappa = 1, I_p = 1 on x < 0."""
    if x >= 0:
        return 2 * np.pi * (expn(3, x) - expn(3, x + L))
    return 2 * np.pi * (expn(3, -x) - expn(3, x + L))


def test_milne_emergent_flux_and_W_converge():
    """Observed on the 16-cell slab: emergent flux at x = +L off by 5.6%,
    1.0%, 0.01% for K = 1, 2, 3; at x = -L 3.6%, 1.5%, 0.9%; the solved
    W (sign flipped, the generator's s convention) within 14%, 10%, 9% of
    the exact flux at the tet centroids, the jump at x = 0 included."""
    L = 1.0
    errs = []
    for K in (1, 2, 3):
        P = ta.problems.milne(8, K)
        ar = P.mesh.arrays()
        op = oracle.OracleProblem(ar, ang_dict(P.ang))
        x, _ = masked_cg(op, reflect_mask(ar, K, P.ang.slm))
        F = op.boundary_flux(x)
        xyz, bnd, kind = ar["coords"], ar["bnd"], ar["face_kind"]
        fx = np.array([xyz[b["p"][:3], 0].mean() for b in bnd])
        right, left = (kind == 0) & (fx > L - 1e-9), (kind == 0) & (fx < -L + 1e-9)
        er = abs(F[right].mean() / _slab_flux(L, L) - 1)
        el = abs(F[left].mean() / -_slab_flux(-L, L) - 1)
        _, T = op.moments(x)
        W = op.flux(T)
        cx = xyz[np.array([t["p"] for t in ar["mesh"]]), 0].mean(axis=1)
        Wex = np.array([_slab_flux(c, L) for c in cx])
        ew = np.abs(-W[0] - Wex).max() / np.abs(Wex).max()
        errs.append((er, el, ew))
    (r1, l1, w1), (r2, l2, w2), (r3, l3, w3) = errs
    assert r1 < 0.07 and r2 < 0.015 and r3 < 0.002
    assert l1 < 0.045 and l2 < 0.02 and l3 < 0.012
    assert w1 < 0.16 and w2 < 0.11 and w3 < 0.10 and w1 > w2 > w3
