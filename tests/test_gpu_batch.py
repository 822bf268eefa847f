"""GPU: batches of independent problems (tr_batch_*).  The union operator
restricted to each problem equals that problem's own operator (oracle, with
the problem's isotropic albedo in massfac[0]) to 1e-10 relative; every
problem of a batched solve reaches its own FP64 residual bar and matches the
problem solved alone."""
import numpy as np
import pytest

import oracle
import transfer_amd as ta
from conftest import ang_dict

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def small_batch():
    return ta.problems.batch(count=6, nrange=(3, 6), seed=11)


def test_batch_operator_matches_each_problem():
    B = small_batch()
    print("[gpu-batch] operator", flush=True)
    ctx = ta.Batch(B.meshes, B.ang, B.albedo)
    aslm = ctx.aslm
    rng = np.random.default_rng(3)
    f = np.zeros((ctx.nP, aslm))
    f[:, :ctx.slm] = rng.standard_normal((ctx.nP, ctx.slm))
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    schedules = [None]
    try:
        ctx.set_schedule(0, J=0, WPB=4, variant=4)
        schedules.append("mfma")
    except ta.TransferError:
        pass
    for sch in schedules:
        if sch is None:
            ctx.set_schedule(0, J=3, WPB=4, variant=3)
        else:
            ctx.set_schedule(0, J=0, WPB=4, variant=4)
        ctx.computeLhs(d_f, d_r)
        r = ctx.copyToHost(d_r).reshape(ctx.nP, aslm)
        for p, m in enumerate(B.meshes):
            a = ang_dict(B.ang)
            mf = np.ones(ctx.slm)
            mf[0] = 1.0 - B.albedo[p]
            a["massfac"] = mf
            op = oracle.OracleProblem(m.arrays(), a)
            lo, hi = int(ctx.voff[p]), int(ctx.voff[p + 1])
            want = op.lhs(f[lo:hi].ravel()).reshape(hi - lo, aslm)
            e = rel(r[lo:hi], want)
            print("[gpu-batch] %s problem %d rel err %.2e" % (sch or "v3", p, e), flush=True)
            assert e <= 1e-10
    ctx.deviceFree(d_f)
    ctx.deviceFree(d_r)


@pytest.mark.parametrize("solver,inner_tol,replace", [(ta.TR_SOLVER_PCG_FP64, 1e-4, True),
                                                      (ta.TR_SOLVER_MIXED, 1e-4, True),
                                                      (ta.TR_SOLVER_MIXED, 1e-2, True),
                                                      (ta.TR_SOLVER_MIXED, 1e-2, False)])
@pytest.mark.parametrize("prec", [2, 4])
def test_batch_solve_matches_individual_solves(solver, inner_tol, replace, prec):
    """Every problem of a batched solve, with Jacobi (2) or the Chebyshev
    polynomial over the union (4), reaches the FP64 residual and the
    solution of the same problem solved alone; the mixed solver with its
    refinement steps continuing each problem's inner CG (residual
    replacement, per problem) and restarting it."""
    B = small_batch()
    print("[gpu-batch] solve solver=%d prec=%d inner_tol=%g replacement=%d" % (solver, prec, inner_tol, replace),
          flush=True)
    ctx = ta.Batch(B.meshes, B.ang, B.albedo)
    ctx.autotune(0)
    if solver == ta.TR_SOLVER_MIXED:
        ctx.autotune(1)
    ctx.set_residual_replacement(replace)
    d_x = ctx.allocVector()
    st, per = ctx.solve_batch(d_x, solver=solver, use_prec=prec, tol_rel=1e-8, inner_tol=inner_tol)
    assert st.status == 0
    x = ctx.copyToHost(d_x).reshape(ctx.nP, ctx.aslm)
    for p, m in enumerate(B.meshes):
        assert per[p]["status"] == 0 and per[p]["true_rel"] <= 1.01e-8
        ang = ta.Angular(2, B.albedo[p], 0.0, 0)
        one = ta.HipContext(m, ang)
        d_y = one.allocVector()
        s1, _ = one.solve(d_y, solver=ta.TR_SOLVER_PCG_FP64, use_prec=2, maxit=20000, tol_abs=0.0, tol_rel=1e-11)
        y = one.copyToHost(d_y).reshape(one.nP, one.aslm)
        lo, hi = int(ctx.voff[p]), int(ctx.voff[p + 1])
        e = rel(x[lo:hi], y)
        print("[gpu-batch] problem %d iters %d inner %d true_rel %.2e vs alone %.2e" % (
            p, per[p]["iters"], per[p]["inner_iters"], per[p]["true_rel"], e), flush=True)
        assert e <= 1e-6
        one.deviceFree(d_y)
    ctx.deviceFree(d_x)
