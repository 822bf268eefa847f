"""The reference's soltype-1 check (main.cpp:190-197: the computed U against
a sphere solution U1(r)), made here with the exact transport solution of the
same two-region sphere (tests/sphere_exact.py: core r < 0.3 with kappa 11
and I_p 1, shell to r = 1 with kappa 1 and no source, no scattering) in
place of the reference's fitted curves, on a curved sphere mesh whose
boundary faces all have normals off the axes (problems.sphere).

With general normals (TR_CTX_GENERAL_NORMALS) the vacuum boundary term is
exact for every face, and the error left is the P_N truncation and the mesh
(the core's staircase edge).  In the optically thin shell the radiation
streams out of a small source, so P_N converges slowly there: measured on
this mesh, 0.71, 0.26, 0.19 and 0.18 at K = 1, 3, 4, 5, with the core at 1%.  The reference's axis rule applies a wrong boundary term on such a
mesh: its shell error is 1.00 at K = 1 and 0.47 at K = 3.  Errors are
relative RMS over the vertices of each region."""
import numpy as np
import pytest

import sphere_exact as se
import transfer_amd as ta

pytestmark = pytest.mark.gpu


def solve_U(P, general):
    ctx = ta.HipContext(P.mesh, P.ang, general_normals=general)
    d = ctx.allocVector()
    st, _ = ctx.solve(d, solver=ta.TR_SOLVER_PCG_FP64, use_prec=2, maxit=40000, tol_abs=0.0, tol_rel=1e-9)
    assert st.status == 0 and st.true_rel <= 1.01e-9
    U, _ = ctx.moments(d)
    ctx.deviceFree(d)
    ctx.close()
    return U


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_sphere_against_exact_transport():
    rs = np.linspace(0.0, 1.0, 401)
    table = se.U_table(rs)
    err = {}
    for K in (1, 3, 4, 5):
        P = ta.problems.sphere(40, maxk=K)
        r = np.linalg.norm(P.mesh.geometry()[0], axis=1)
        Ue = np.interp(r, rs, table)
        core, shell = r < 0.2, r > 0.45
        for general in ((True, False) if K < 4 else (True,)):
            print("[gpu-sphere] solving K=%d general=%s (%d vertices)" % (K, general, len(r)), flush=True)
            U = solve_U(P, general)
            assert np.all(np.isfinite(U))
            e = (rel_rms(U[core], Ue[core]), rel_rms(U[shell], Ue[shell]), rel_rms(U, Ue))
            err[(K, general)] = e
            print("[gpu-sphere] K=%d general=%-5s rel RMS error: core %.3f shell %.3f all %.3f" % ((K, general) + e),
                  flush=True)
    # the exact boundary term beats the axis rule in the shell, the shell
    # error falls with the angular order, and the core is right to 3%
    for K in (1, 3):
        assert err[(K, True)][1] < 0.8 * err[(K, False)][1]
    assert err[(5, True)][1] < err[(3, True)][1] < err[(1, True)][1]
    assert err[(4, True)][1] < err[(1, True)][1]
    assert max(err[(K, True)][0] for K in (1, 3, 4, 5)) <= 0.03
