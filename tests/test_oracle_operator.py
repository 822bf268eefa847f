"""Pins the oracle's operator (its restatement of kernels.cu) against a dense
weak-form assembly from coordinates, and checks the invariants the reference
relies on (symmetry, SPD, zero row sums of the stiffness, RHS closed form)."""
import numpy as np
import pytest

import oracle
import transfer_amd as ta
from angular_np import couplings
from brute import assemble
from conftest import ang_dict


def tiny(n, jitter, K, albedo=0.0, g=0.0, L=0, seed=7):
    m = ta.Mesh.box((n, n, n), (0, 0, 0), (1, 1.5, 2), jitter=jitter, seed=seed)
    rng = np.random.default_rng(seed)
    kappa = 0.5 + rng.random(m.nT)
    Ip = rng.random(m.nT)
    m.set_fields(kappa, Ip)
    a = ta.Angular(K, albedo, g, L)
    return m, a, kappa, Ip


def dense_oracle(op, diag=False):
    s, a = op.slm, op.aslm
    N = op.nP * s
    A = np.zeros((N, N))
    for col in range(N):
        f = np.zeros(op.N())
        v, lm = divmod(col, s)
        f[v * a + lm] = 1
        r = (op.lhs_diag(f) if diag else op.lhs(f)).reshape(op.nP, a)[:, :s]
        A[:, col] = r.reshape(-1)
    return A


@pytest.mark.parametrize("n,K,albedo,g,L", [(1, 1, 0.0, 0.0, 0), (2, 2, 0.0, 0.0, 0),
                                             (2, 1, 0.5, 0.0, 0), (2, 2, 0.9, 0.6, 7)])
def test_operator_matches_weak_form(n, K, albedo, g, L):
    m, a, kappa, Ip = tiny(n, 0.15, K, albedo, g, L)
    op = oracle.OracleProblem(m.arrays(), oracle.angular(K, albedo, g, L))
    coords, tets, _, _ = m.geometry()
    E, O, mf = couplings(K, albedo, g, L)
    for diag in (False, True):
        Ab, bb = assemble(coords, tets, kappa, Ip, E, O, mf, diag=diag)
        Ao = dense_oracle(op, diag=diag)
        np.testing.assert_allclose(Ao, Ab, rtol=0, atol=1e-12 * np.abs(Ab).max())
    b = op.rhs().reshape(op.nP, op.aslm)[:, :op.slm].reshape(-1)
    np.testing.assert_allclose(b, bb, atol=1e-14)


def test_operator_symmetric_positive_definite():
    m, a, kappa, Ip = tiny(2, 0.15, 2)
    op = oracle.OracleProblem(m.arrays(), oracle.angular(2))
    A = dense_oracle(op)
    # the reference's own check (main.cpp:106-113), at 100 ulp; entries that
    # are pure cancellation residue (~1e-21) get an absolute floor
    floor = 1e-15 * np.abs(A).max()
    assert np.all(np.abs(A - A.T) <= 100 * 2.2e-16 * (np.abs(A) + np.abs(A.T)) + floor)
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    D = dense_oracle(op, diag=True)
    assert np.allclose(D, D.T, atol=1e-14 * np.abs(D).max())
    assert np.linalg.eigvalsh(0.5 * (D + D.T)).min() > 0


def test_stiffness_rows_sum_to_zero_and_constant_field():
    m, a, kappa, Ip = tiny(3, 0.15, 2, albedo=0.3)
    arr = m.arrays()
    op = oracle.OracleProblem(arr, oracle.angular(2, 0.3))
    rng = np.random.default_rng(1)
    f = np.zeros((op.nP, op.aslm))
    f[:, :op.slm] = rng.standard_normal((op.nP, op.slm))
    r = op.volume(f.ravel()).reshape(op.nP, op.aslm)
    # sum_i of the volume rows = mass only: sum_t kappa V massfac / 4 sum_k f_k
    mesh = arr["mesh"]
    want = np.zeros(op.slm)
    for t in range(len(mesh)):
        want += mesh["kappa_volume"][t] / 4 * f[mesh["p"][t], :op.slm].sum(0)
    want *= oracle.angular(2, 0.3)["massfac"]
    np.testing.assert_allclose(r[:, :op.slm].sum(0), want, rtol=1e-12, atol=1e-12)
    # constant f_00 = 1: stiffness vanishes, mass gives sum kappa V/4 massfac_0
    f = np.zeros((op.nP, op.aslm))
    f[:, 0] = 1
    r = op.volume(f.ravel()).reshape(op.nP, op.aslm)
    kv4 = np.zeros(op.nP)
    for t in range(len(mesh)):
        kv4[mesh["p"][t]] += mesh["kappa_volume"][t] / 4
    np.testing.assert_allclose(r[:, 0], 0.7 * kv4, rtol=1e-12)
    assert np.abs(r[:, 1:]).max() < 1e-12 * np.abs(r[:, 0]).max()


@pytest.mark.parametrize("use_prec", [False, True])
def test_pcg_solves_the_system(use_prec):
    m, a, kappa, Ip = tiny(2, 0.1, 2)
    op = oracle.OracleProblem(m.arrays(), oracle.angular(2))
    f, st, hist = op.pcg(use_prec=use_prec, tol_abs=0, tol_rel=1e-13, maxit=2000, inner_tol=1e-12,
                         record=3)
    A = dense_oracle(op)
    b = op.rhs().reshape(op.nP, op.aslm)[:, :op.slm].reshape(-1)
    x = np.linalg.solve(A, b)
    fx = f.reshape(op.nP, op.aslm)[:, :op.slm].reshape(-1)
    np.testing.assert_allclose(fx, x, rtol=1e-9, atol=1e-10 * np.abs(x).max())
    assert st.rnorm <= 1e-13 * st.bnorm
    assert hist.shape == (3, 3) and np.all(hist[:, 0] > 0)
