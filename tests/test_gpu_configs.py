"""GPU parity at the benchmark configurations' full sizes, against the
oracle on the same inputs:

- the 96^3 Henyey-Greenstein ball (2,780,400 tets, K = 5, slm 66): the
  FP64 operator at 1e-10 relative to max|A f| (the default schedule), the
  exact FP32 operator at 1e-5, the BF16 matrix-core one (hi/lo split) at
  2e-4, and mixed solves to a 1e-8 FP64 residual with each FP32 operator
  that the oracle re-checks on the host;
- the 256-slab batch (union of 12.1M vertices, K = 2): the union operator
  restricted to 8 sampled problems (the smallest, the largest and six
  spread between) against each problem's own oracle at 1e-10, and the
  batched mixed solve of all 256 problems to 1e-8 (each problem's FP64
  residual on the device; the oracle re-checks two of them);
- the 32^3 blob cube (196,608 tets, K = 3) of the learned part: the learned
  guess (network trained on smaller solver-generated cubes) in front of the
  mixed solve, the oracle re-checking the solution's residual.

Each case prints its sizes and errors before it asserts."""
import numpy as np
import pytest

import oracle
import transfer_amd as ta
from conftest import ang_dict

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rvec(nP, slm, aslm, seed):
    f = np.zeros((nP, aslm))
    f[:, :slm] = np.random.default_rng(seed).standard_normal((nP, slm))
    return f.reshape(-1)


def oracle_residual(op, f):
    b = op.rhs()
    return op.norm(b - op.lhs(f)) / op.norm(b)


def test_ball96_hg_operator_and_mixed_solve():
    P = ta.problems.ball(96, maxk=5, kind="hg")
    print("[gpu-configs] ball96-hg: nP=%d nT=%d slm=%d" % (P.mesh.nP, P.mesh.nT, P.ang.slm), flush=True)
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    f = rvec(op.nP, op.slm, op.aslm, 1)
    want = op.lhs(f)
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    ctx.computeLhs(d_f, d_r)
    e64 = rel(ctx.copyToHost(d_r), want)
    print("[gpu-configs] ball96-hg FP64 %s: rel err %.2e" % (ctx.kernel_name(0), e64), flush=True)
    assert e64 <= 1e-10
    # the FP64 operator on the FP64 matrix cores (parity-class blocks), and
    # the plain FP64 solve with it, which the oracle re-checks
    default = ctx.schedule()
    ctx.set_schedule(0, J=0, WPB=8, variant=5)
    ctx.computeLhs(d_f, d_r)
    e5 = rel(ctx.copyToHost(d_r), want)
    print("[gpu-configs] ball96-hg FP64 %s: rel err %.2e" % (ctx.kernel_name(0), e5), flush=True)
    assert e5 <= 1e-10
    d_x = ctx.allocVector()
    st, _ = ctx.solve(d_x, solver=ta.TR_SOLVER_PCG_FP64, use_prec=0, maxit=5000, tol_abs=0.0, tol_rel=1e-8)
    res = oracle_residual(op, ctx.copyToHost(d_x))
    print("[gpu-configs] ball96-hg FP64 PCG on %s: %d iterations, %.3f s, true rel %.2e (oracle %.2e)" % (
        ctx.kernel_name(0), st.iters, st.seconds, st.true_rel, res), flush=True)
    assert st.status == 0 and st.true_rel <= 1e-8 and res <= 1.5e-8
    ctx.deviceFree(d_x)
    ctx.set_schedule(0, J=default["J"], WPB=default["WPB"], variant=default["variant"])
    errs = {}
    for prec, bar in ((0, 1e-5), (3, 2e-4), (4, 1e-6)):
        ctx.set_angular_precision(prec)
        ctx.computeLhsPrec(1, d_f, d_r)
        errs[prec] = rel(ctx.copyToHost(d_r), want)
        print("[gpu-configs] ball96-hg FP32 angular precision %d %s: rel err %.2e" % (
            prec, ctx.kernel_name(1), errs[prec]), flush=True)
        assert errs[prec] <= bar
    for prec in (0, 3):
        ctx.set_angular_precision(prec)
        d_x = ctx.allocVector()
        st, _ = ctx.solve(d_x, solver=ta.TR_SOLVER_MIXED, use_prec=2, maxit=30, tol_abs=0.0, tol_rel=1e-8,
                          inner_maxit=20000, inner_tol=1e-5)
        res = oracle_residual(op, ctx.copyToHost(d_x))
        print("[gpu-configs] ball96-hg mixed solve (angular precision %d): %d refinements, %d inner, %.3f s, "
              "true rel %.2e (oracle %.2e)" % (prec, st.iters, st.preck, st.seconds, st.true_rel, res), flush=True)
        assert st.status == 0 and st.true_rel <= 1e-8 and res <= 1.5e-8
        ctx.deviceFree(d_x)
    # residual replacement against the restarted refinement, both replacing
    # (or restarting) every 1e-2 of inner reduction: observed 686 against
    # 763 inner iterations
    inner = {}
    for replace in (True, False):
        ctx.set_residual_replacement(replace)
        d_x = ctx.allocVector()
        st, _ = ctx.solve(d_x, solver=ta.TR_SOLVER_MIXED, use_prec=0, maxit=30, tol_abs=0.0, tol_rel=1e-8,
                          inner_maxit=20000, inner_tol=1e-2)
        res = oracle_residual(op, ctx.copyToHost(d_x))
        print("[gpu-configs] ball96-hg mixed solve, replacement %d: %d steps, %d inner, %.3f s, true rel %.2e "
              "(oracle %.2e)" % (replace, st.iters, st.preck, st.seconds, st.true_rel, res), flush=True)
        assert st.status == 0 and st.fallback == 0 and st.true_rel <= 1e-8 and res <= 1.5e-8
        inner[replace] = st.preck
        ctx.deviceFree(d_x)
    assert inner[True] <= 0.95 * inner[False]
    ctx.close()


def test_batch256_union_operator_and_batched_solve():
    B = ta.problems.batch(256)
    ctx = ta.Batch(B.meshes, B.ang, B.albedo)
    nP, slm, aslm = ctx.nP, ctx.slm, ctx.aslm
    sizes = np.array([m.nP for m in B.meshes])
    print("[gpu-configs] batch256: union nP=%d, problems %d..%d vertices" % (nP, sizes.min(), sizes.max()),
          flush=True)
    f = rvec(nP, slm, aslm, 3).reshape(nP, aslm)
    d_f, d_r = ctx.vector_from(f.ravel()), ctx.allocVector()
    ctx.computeLhs(d_f, d_r)
    r = ctx.copyToHost(d_r).reshape(nP, aslm)
    ctx.deviceFree(d_f)
    order = np.argsort(sizes, kind="stable")
    sample = sorted(set(int(order[i]) for i in np.linspace(0, len(order) - 1, 8).round().astype(int)))
    for p in sample:
        lo, hi = int(ctx.voff[p]), int(ctx.voff[p + 1])
        op = oracle.OracleProblem(B.meshes[p].arrays(), ang_dict(ta.Angular(B.ang.maxk, B.albedo[p])))
        e = rel(r[lo:hi].ravel(), op.lhs(f[lo:hi].ravel()))
        print("[gpu-configs] batch256 problem %d (%d vertices, albedo %.3f): union operator rel err %.2e" % (
            p, hi - lo, B.albedo[p], e), flush=True)
        assert e <= 1e-10
    d_x = ctx.allocVector()
    st, per = ctx.solve_batch(d_x, solver=ta.TR_SOLVER_MIXED, use_prec=2, maxit=30, tol_rel=1e-8)
    worst = max(q["true_rel"] for q in per)
    print("[gpu-configs] batch256 mixed solve: %.2f s, %d/%d converged, worst true rel %.2e" % (
        st.seconds, sum(q["status"] == 0 for q in per), len(per), worst), flush=True)
    assert st.status == 0 and all(q["status"] == 0 for q in per) and worst <= 1e-8
    x = ctx.copyToHost(d_x).reshape(nP, aslm)
    for p in (sample[0], sample[-1]):
        lo, hi = int(ctx.voff[p]), int(ctx.voff[p + 1])
        op = oracle.OracleProblem(B.meshes[p].arrays(), ang_dict(ta.Angular(B.ang.maxk, B.albedo[p])))
        res = oracle_residual(op, x[lo:hi].ravel())
        print("[gpu-configs] batch256 problem %d: oracle residual %.2e" % (p, res), flush=True)
        assert res <= 1.5e-8
    ctx.close()


def test_cube32_learned_guess_and_mixed_solve():
    from transfer_amd import learned
    feat, targ, _ = learned.generate(list(range(100, 124)), n=10, maxk=3, chunk=24)
    g = learned.LearnedGuess(seed=2)
    g.fit(feat, targ, steps=600, batch=8192, lr=3e-3)
    P = ta.problems.cube(32, maxk=3, seed=900)
    print("[gpu-configs] cube32: nP=%d nT=%d albedo %.3f" % (P.mesh.nP, P.mesh.nT, P.albedo), flush=True)
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    kw = dict(solver=ta.TR_SOLVER_MIXED, use_prec=2, maxit=30, tol_abs=0.0, tol_rel=1e-8, inner_maxit=20000,
              inner_tol=1e-5)
    d0 = ctx.allocVector()
    s0, _ = ctx.solve(d0, **kw)
    d1 = g.write_guess(ctx, P.mesh, P.albedo)
    s1, _ = ctx.solve(d1, x0=True, **kw)
    res = oracle_residual(op, ctx.copyToHost(d1))
    print("[gpu-configs] cube32 mixed solve: zero start %d inner %.4f s; learned start %d inner %.4f s, "
          "fallback %d, oracle residual %.2e" % (s0.preck, s0.seconds, s1.preck, s1.seconds, s1.fallback, res),
          flush=True)
    assert s0.status == 0 and s1.status == 0 and res <= 1.5e-8
    ctx.close()
