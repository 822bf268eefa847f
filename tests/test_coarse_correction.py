"""What the scalar radiation can and cannot do for the solver (CPU, the
assembled operator in scipy).

The learned network predicts the scalar radiation U = 4 pi f_00.  Used as a
coarse correction, that prediction spans a coarse space of lm = 0 functions
on vertex aggregates (runs of consecutive Morton-ordered vertices) weighted
by the predicted profile; used as a start, it sets f_00 and leaves the
higher moments at zero.  This test measures both with the best prediction
any network could give, the exact solution's own U, against Jacobi-PCG on a
scattering cube (albedo 0.96, K = 3):

- the scalar coarse space (additive two-level, exact coarse solve) saves
  under 10% of the iterations, with the exact profile as with a constant
  one, while the coarse space of all harmonics on the same aggregates
  (P (x) I) saves over 25%: the slow modes of the even-parity system are
  not scalar;
- the exact f_00 as the only nonzero part of the start leaves a residual
  above ||b|| (the higher moments' rows), so the iteration count barely
  changes.

These numbers are why the library uses the prediction as the initial guess
only (with the residual fallback) and accelerates the solve with a
polynomial (Chebyshev) preconditioner instead."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import transfer_amd as ta


def assembled(P):
    st = ta.Stencil(P.mesh.view, P.ang.view)
    nP, slm = st.nP, st.slm

    def ell(nbr, coef, rows):
        return sp.csr_matrix((coef.ravel(), (np.tile(rows, nbr.shape[0]), nbr.ravel())), shape=(nP, nP))

    Ms = [ell(st.nbr, st.coef[:, c, :], np.arange(nP)) for c in range(st.NC)]
    Ms += [ell(st.bnbr, st.bcoef[:, n, :], st.bvert) for n in range(3)]
    Oms = [np.diag(st.massfac)] + [st.omc[c] for c in range(st.NC - 1)] + [st.onc[n] for n in range(3)]
    b = np.zeros((nP, slm))
    b[:, 0] = st.rhs
    return Ms, Oms, b.ravel(), nP, slm


def apply(Ms, Oms, f, n, slm):
    F = f.reshape(n, slm)
    R = np.zeros_like(F)
    for M, O in zip(Ms, Oms):
        R += (M @ F) @ O.T
    return R.ravel()


def pcg(A, b, M, x0=None, tol=1e-8, maxit=5000):
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A(x)
    r0 = np.linalg.norm(r) / np.linalg.norm(b)
    z = M(r)
    p, rz, nb = z.copy(), r @ z, np.linalg.norm(b)
    for it in range(maxit):
        Ap = A(p)
        al = rz / (p @ Ap)
        x += al * p
        r -= al * Ap
        if np.linalg.norm(r) <= tol * nb:
            return it + 1, x, r0
        z = M(r)
        rz2 = r @ z
        p = z + rz2 / rz * p
        rz = rz2
    return maxit, x, r0


@pytest.fixture(scope="module")
def cube():
    P = ta.problems.cube(10, maxk=3, seed=10)
    assert P.albedo > 0.9
    Ms, Oms, b, nP, slm = assembled(P)
    A = lambda f: apply(Ms, Oms, f, nP, slm)
    d = sum(np.outer(M.diagonal(), np.diag(O)) for M, O in zip(Ms, Oms)).ravel()
    n_j, x, _ = pcg(A, b, lambda r: r / d)
    return dict(Ms=Ms, Oms=Oms, b=b, nP=nP, slm=slm, A=A, d=d, n_j=n_j, x=x)


def test_scalar_coarse_space_does_not_accelerate(cube):
    Ms, Oms, nP, slm, A, d, b = cube["Ms"], cube["Oms"], cube["nP"], cube["slm"], cube["A"], cube["d"], cube["b"]
    U = cube["x"].reshape(nP, slm)[:, 0]
    size = 16
    owner = np.arange(nP) // size
    nC = owner[-1] + 1
    res = {}
    for name, w in (("constant", np.ones(nP)), ("exact U", U / U.mean())):
        Z = sp.csr_matrix((w, (np.arange(nP), owner)), shape=(nP, nC))
        Ac = sum(O[0, 0] * (Z.T @ M @ Z) for M, O in zip(Ms, Oms)).tocsc()
        lu = spl.splu(Ac)

        def prec(r):
            y = np.zeros((nP, slm))
            y[:, 0] = Z @ lu.solve(Z.T @ r.reshape(nP, slm)[:, 0])
            return r / d + y.ravel()

        res[name] = pcg(A, b, prec)[0]
    # all harmonics on the same aggregates: P (x) I, coarse solve by CG to 1e-10
    P1 = sp.csr_matrix((np.ones(nP), (np.arange(nP), owner)), shape=(nP, nC))
    Mc = [(P1.T @ M @ P1).tocsr() for M in Ms]
    dc = sum(np.outer(M.diagonal(), np.diag(O)) for M, O in zip(Mc, Oms)).ravel()
    Ac = lambda f: apply(Mc, Oms, f, nC, slm)

    def prec_full(r):
        rc = (P1.T @ r.reshape(nP, slm)).ravel()
        yc = pcg(Ac, rc, lambda q: q / dc, tol=1e-10)[1]
        return r / d + (P1 @ yc.reshape(nC, slm)).ravel()

    res["all harmonics"] = pcg(A, b, prec_full)[0]
    print("[coarse] Jacobi-PCG %d iterations; additive two-level with the scalar space: constant %d, exact U %d; "
          "with all harmonics %d" % (cube["n_j"], res["constant"], res["exact U"], res["all harmonics"]))
    assert res["exact U"] >= 0.9 * cube["n_j"] and res["constant"] >= 0.9 * cube["n_j"]
    assert res["all harmonics"] <= 0.75 * cube["n_j"]


def test_exact_scalar_start_leaves_the_higher_moments(cube):
    nP, slm, A, d, b = cube["nP"], cube["slm"], cube["A"], cube["d"], cube["b"]
    x0 = np.zeros((nP, slm))
    x0[:, 0] = cube["x"].reshape(nP, slm)[:, 0]
    n, _, r0 = pcg(A, b, lambda r: r / d, x0=x0.ravel())
    print("[coarse] exact f_00 start: initial residual %.2f ||b||, %d iterations (zero start %d)" % (r0, n, cube["n_j"]))
    assert r0 > 1.0 and n >= 0.9 * cube["n_j"]
