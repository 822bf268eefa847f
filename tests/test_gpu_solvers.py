"""Device-resident CG (FP64) and FP64-refined mixed-precision solves, with
no, Jacobi (2), multilevel (3) or Chebyshev polynomial (4) preconditioning,
against the oracle's FP64 PCG solution on the same inputs.

Bars: the FP64 true residual ||b - A x|| / ||b|| of every returned solution
meets the requested tolerance (checked on the GPU with the FP64 operator and
again on the host with the oracle operator); the solution's distance to the
oracle's converged FP64 solution is reported and bounded by 1e-6 relative."""
import numpy as np
import pytest

import oracle
import transfer_amd as ta
from conftest import ang_dict

pytestmark = pytest.mark.gpu


def build(name):
    if name == "slab8":
        return ta.problems.slab(n=8)
    if name == "ball16":
        return ta.problems.ball(16)
    if name == "cube8":
        return ta.problems.cube(8)
    if name == "ball12-hg":
        return ta.problems.ball(12, maxk=2, kind="hg")
    if name == "ball10-K5-hg":
        return ta.problems.ball(10, maxk=5, kind="hg")
    if name == "ball16-K5":
        return ta.problems.ball(16, maxk=5)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["slab8", "ball16", "cube8", "ball12-hg", "ball10-K5-hg"])
@pytest.mark.parametrize("solver,prec", [(ta.TR_SOLVER_PCG_FP64, 0), (ta.TR_SOLVER_PCG_FP64, 2),
                                         (ta.TR_SOLVER_PCG_FP64, 3), (ta.TR_SOLVER_PCG_FP64, 4),
                                         (ta.TR_SOLVER_MIXED, 2), (ta.TR_SOLVER_MIXED, 3),
                                         (ta.TR_SOLVER_MIXED, 4)])
def test_solver_reaches_fp64_residual(name, solver, prec):
    print("[gpu-solvers]", name, solver, prec, flush=True)
    P = build(name)
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    tol = 1e-8 if solver == ta.TR_SOLVER_MIXED else 1e-10
    d_f = ctx.allocVector()
    st, _ = ctx.solve(d_f, solver=solver, use_prec=prec, maxit=20000 if solver != ta.TR_SOLVER_MIXED else 30,
                      tol_abs=0.0, tol_rel=tol, inner_maxit=20000, inner_tol=1e-4)
    f = ctx.copyToHost(d_f)
    b = op.rhs()
    res = op.norm(b - op.lhs(f)) / op.norm(b)
    ref, st_or, _ = op.pcg(tol_abs=0.0, tol_rel=1e-12, maxit=40000)
    err = np.abs(f - ref).max() / np.abs(ref).max()
    print("[gpu-solvers] %s solver=%d prec=%d iters=%d inner=%d true_rel(gpu)=%.2e (oracle)=%.2e "
          "err_vs_fp64=%.2e %.3fs" % (name, solver, prec, st.iters, st.preck, st.true_rel, res, err, st.seconds),
          flush=True)
    assert st.status == 0
    assert st.true_rel <= tol * 1.01 and res <= tol * 1.5
    assert err <= 1e-6
    ctx.close()


@pytest.mark.parametrize("name", ["ball16", "ball12-hg", "ball10-K5-hg"])
@pytest.mark.parametrize("prec", [0, 2, 4])
@pytest.mark.parametrize("inner_tol", [1e-2, 1e-5])
def test_residual_replacement(name, prec, inner_tol):
    """The mixed solver's refinement steps with residual replacement (the
    inner CG continued across steps, its residual replaced by the true FP64
    one; the default) and restarted: both reach the FP64 residual, checked
    again by the oracle, and the oracle's solution, with no fallback to the
    FP64 solver.  At 1e-5 the inner CG's recursive residual can drift from
    the true one before a replacement; such steps restart."""
    P = build(name)
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    b = op.rhs()
    ref, _, _ = op.pcg(tol_abs=0.0, tol_rel=1e-12, maxit=40000)
    for replace in (True, False):
        ctx.set_residual_replacement(replace)
        d_f = ctx.allocVector()
        st, _ = ctx.solve(d_f, solver=ta.TR_SOLVER_MIXED, use_prec=prec, maxit=30, tol_abs=0.0, tol_rel=1e-8,
                          inner_maxit=20000, inner_tol=inner_tol)
        f = ctx.copyToHost(d_f)
        ctx.deviceFree(d_f)
        res = op.norm(b - op.lhs(f)) / op.norm(b)
        err = np.abs(f - ref).max() / np.abs(ref).max()
        print("[gpu-solvers] %s prec=%d inner_tol=%g replacement=%d: %d steps, %d inner, true_rel %.2e "
              "(oracle %.2e), err_vs_fp64 %.2e, fallback %d" % (name, prec, inner_tol, replace, st.iters, st.preck,
                                                                st.true_rel, res, err, st.fallback), flush=True)
        assert st.status == 0 and st.fallback == 0
        assert st.true_rel <= 1e-8 * 1.01 and res <= 1.5e-8
        assert err <= 1e-6
    ctx.close()


@pytest.mark.parametrize("name", ["ball10-K5-hg", "ball16-K5"])
@pytest.mark.parametrize("prec", [0, 2, 4])
def test_fp32_matrix_core_operator_solves(name, prec):
    """The mixed solver with the FP32 matrix-core operator (angular precision
    4: k_apply5s with FP32 products), its fused p.Ap dot (no
    preconditioner, Jacobi) and its fused Chebyshev steps, against the
    oracle: the FP64 residual reached and re-checked by the oracle, the
    oracle's solution to 1e-6, no fallback."""
    P = build(name)
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    ctx.set_angular_precision(4)
    d_f = ctx.allocVector()
    st, _ = ctx.solve(d_f, solver=ta.TR_SOLVER_MIXED, use_prec=prec, maxit=30, tol_abs=0.0, tol_rel=1e-8,
                      inner_maxit=20000, inner_tol=1e-2 if prec != 4 else 1e-4)
    kname = ctx.kernel_name(1)
    f = ctx.copyToHost(d_f)
    b = op.rhs()
    res = op.norm(b - op.lhs(f)) / op.norm(b)
    ref, _, _ = op.pcg(tol_abs=0.0, tol_rel=1e-12, maxit=40000)
    err = np.abs(f - ref).max() / np.abs(ref).max()
    print("[gpu-solvers] %s FP32 matrix-core %s prec=%d: %d steps, %d inner, true_rel %.2e (oracle %.2e), "
          "err_vs_fp64 %.2e" % (name, kname, prec, st.iters, st.preck, st.true_rel, res, err), flush=True)
    assert kname.endswith(", 4, 0>"), kname
    assert st.status == 0 and st.fallback == 0
    assert st.true_rel <= 1e-8 * 1.01 and res <= 1.5e-8
    assert err <= 1e-6
    ctx.close()


def test_multilevel_hierarchy_shapes():
    """Every hierarchy shape gives a converged solve with the same solution
    (the preconditioner changes the path, not the answer)."""
    P = ta.problems.cube(10, seed=2)
    ctx = ta.HipContext(P.mesh, P.ang)
    ref = None
    for coarsest, agg, nu in ((64, 8, 2), (500, 16, 4), (100000, 8, 2)):
        ctx.set_multilevel(coarsest, agg, nu)
        d = ctx.allocVector()
        st, _ = ctx.solve(d, solver=ta.TR_SOLVER_PCG_FP64, use_prec=3, maxit=5000, tol_abs=0.0, tol_rel=1e-10)
        x = ctx.copyToHost(d)
        ctx.deviceFree(d)
        print("[gpu-solvers] multilevel %d/%d/%d: %d iters" % (coarsest, agg, nu, st.iters), flush=True)
        assert st.status == 0 and st.true_rel <= 1.01e-10
        if ref is None:
            ref = x
        assert np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max()


@pytest.mark.parametrize("case", ["fused", "fused-valu", "unfused", "reflective"])
def test_chebyshev_preconditioner(case):
    """The Chebyshev polynomial preconditioner at several degrees, with its
    steps fused into the matrix-core operator's write ("fused") or into the
    buffer-load VALU schedule's ("fused-valu"), or run as operator +
    elementwise pass on the plain VALU schedule ("unfused"), and on a
    problem with reflecting planes (vanishing entries): every solve reaches
    the FP64 residual, higher degrees take fewer outer iterations, and all
    return the same solution."""
    if case == "reflective":
        P = ta.problems.milne(n=12, maxk=2)
    else:
        P = ta.problems.cube(10, seed=4)
    ctx = ta.HipContext(P.mesh, P.ang)
    if case in ("fused-valu", "unfused"):
        variant = 3 if case == "fused-valu" else 1
        ctx.set_schedule(0, J=4, WPB=4, variant=variant)
        ctx.set_schedule(1, J=4, WPB=4, variant=variant)
    ref, its = None, []
    for solver in (ta.TR_SOLVER_PCG_FP64, ta.TR_SOLVER_MIXED):
        for deg in (1, 3, 6):
            ctx.set_chebyshev(deg, 100.0, 2 if deg < 6 else 0)
            d = ctx.allocVector()
            st, _ = ctx.solve(d, solver=solver, use_prec=4, maxit=20000 if solver != ta.TR_SOLVER_MIXED else 30,
                              tol_abs=0.0, tol_rel=1e-10, inner_maxit=20000, inner_tol=1e-4)
            x = ctx.copyToHost(d)
            ctx.deviceFree(d)
            n_it = st.iters if solver == ta.TR_SOLVER_PCG_FP64 else st.preck
            print("[gpu-solvers] chebyshev %s solver=%d degree %d: %d iterations, true rel %.2e" % (
                case, solver, deg, n_it, st.true_rel), flush=True)
            assert st.status == 0 and st.true_rel <= 1.01e-10
            if solver == ta.TR_SOLVER_PCG_FP64:
                its.append(n_it)
            if ref is None:
                ref = x
            assert np.abs(x - ref).max() <= 1e-7 * np.abs(ref).max()
    assert its[0] > its[1] > its[2]


# ---------------------------------------------------------------------------
# Failure handling: every condition is forced from the host (bad starting
# vectors, iteration caps, a refinement threshold no step can meet), so each
# kernel still runs to its normal end.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("solver", [ta.TR_SOLVER_PCG_FP64, ta.TR_SOLVER_MIXED])
@pytest.mark.parametrize("guess", ["nan", "inf", "negative", "huge"])
def test_bad_guess_falls_back_to_zero_start(solver, guess):
    """A NaN, Inf or adversarial x0 (-1e12 in every lm = 0 entry: x0.b < 0,
    so no positive multiple of it is a start) is rejected: the solve starts
    from 0 (fallback 1) and converges to the FP64 tolerance.  A huge random
    x0 (1e12 N(0,1)) is scaled to its energy-optimal multiple first (or
    rejected) and the solve converges either way."""
    P = build("ball12-hg")
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    rng = np.random.default_rng(3)
    x0 = np.zeros((op.nP, op.aslm))
    x0[:, :op.slm] = {"nan": np.nan, "inf": np.inf}.get(guess, 0.0)
    if guess == "huge":
        x0[:, :op.slm] = 1e12 * rng.standard_normal((op.nP, op.slm))
    if guess == "negative":
        x0[:, 0] = -1e12
    d_f = ctx.vector_from(x0.ravel())
    st, _ = ctx.solve(d_f, solver=solver, use_prec=2, x0=True, maxit=20000 if solver != ta.TR_SOLVER_MIXED else 30,
                      tol_abs=0.0, tol_rel=1e-8, inner_maxit=20000, inner_tol=1e-4)
    f = ctx.copyToHost(d_f)
    b = op.rhs()
    res = op.norm(b - op.lhs(f)) / op.norm(b)
    print("[gpu-solvers] guess=%s solver=%d fallback=%d status=%d true_rel=%.2e" % (guess, solver, st.fallback,
                                                                                     st.status, res), flush=True)
    assert st.status == 0 and np.isfinite(f).all() and res <= 1.5e-8
    assert st.fallback == 1 or guess == "huge"
    ctx.close()


@pytest.mark.parametrize("solver", [ta.TR_SOLVER_REFERENCE, ta.TR_SOLVER_PCG_FP64, ta.TR_SOLVER_MIXED])
def test_iteration_cap_reports_noconv(solver):
    """maxit = 1 (main.cpp:142's cap, tightened) ends without convergence:
    status TR_ERR_NOCONV, a finite iterate."""
    P = build("ball16")
    ctx = ta.HipContext(P.mesh, P.ang)
    d_f = ctx.allocVector()
    st, _ = ctx.solve(d_f, solver=solver, use_prec=2 if solver != ta.TR_SOLVER_REFERENCE else 0, maxit=1,
                      tol_abs=0.0, tol_rel=1e-10, inner_maxit=5, inner_tol=1e-8)
    f = ctx.copyToHost(d_f)
    print("[gpu-solvers] maxit=1 solver=%d status=%d iters=%d" % (solver, st.status, st.iters), flush=True)
    assert st.status == ta.TR_ERR_NOCONV and st.iters <= 1 and np.isfinite(f).all()
    ctx.close()


@pytest.mark.parametrize("prec", [2, 4])
def test_mixed_stagnation_falls_back_to_fp64(prec):
    """A refinement threshold no step can meet (each step must cut the true
    residual 1e30-fold) sends the mixed solve to the plain FP64 solver after
    its first step: fallback 2, converged, the oracle's residual re-checked.
    An inner stagnation window of 2 iterations (CG must cut ||r|| by 10%
    every 2 iterations) stops inner solves early, and refinement still
    converges."""
    P = build("ball12-hg")
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    b = op.rhs()
    for stag, window, want_fb in ((1e-30, 100, 2), (0.5, 2, 0)):
        ctx.set_refinement(stag, window)
        d_f = ctx.allocVector()
        st, _ = ctx.solve(d_f, solver=ta.TR_SOLVER_MIXED, use_prec=prec, maxit=60, tol_abs=0.0, tol_rel=1e-8,
                          inner_maxit=20000, inner_tol=1e-4)
        f = ctx.copyToHost(d_f)
        res = op.norm(b - op.lhs(f)) / op.norm(b)
        print("[gpu-solvers] stagnation=%g window=%d prec=%d: fallback=%d status=%d iters=%d inner=%d "
              "true_rel=%.2e" % (stag, window, prec, st.fallback, st.status, st.iters, st.preck, res), flush=True)
        assert st.status == 0 and res <= 1.5e-8
        if want_fb == 2:
            assert st.fallback == 2
        ctx.deviceFree(d_f)
    ctx.close()


def test_zero_opacity_is_rejected():
    """kappa = 0 in one tet: the stiffness term's 1/(9 kappa V)
    (kernels.cu:252) has no value; context creation fails with TR_ERR_ARG."""
    m = ta.Mesh.box((3, 3, 3), (0, 0, 0), (1, 1, 1))
    ar = m.arrays()
    ar["mesh"]["kappa_volume"][5] = 0.0
    with pytest.raises(ta.TransferError) as e:
        ta.HipContext(m, ta.Angular(1), mesh_view=ta.mesh_view_of(ar))
    assert e.value.code == ta.TR_ERR_ARG


def test_operator_timing_modes():
    """tr_bench_operator's modes: the FP64 and FP32 operators (0, 1), the
    FP64 kernel alone (2), the FP32 operator with its fused dot (3) and that
    after a 1 GiB copy (4) all run and report positive per-application
    times, and a bad mode is an argument error."""
    P = build("ball16-K5")
    ctx = ta.HipContext(P.mesh, P.ang)
    ctx.set_angular_precision(3)
    for mode in range(5):
        w, k, kmin, kmax = ctx.bench_operator(1, 3, precision=mode)
        print("[gpu-solvers] bench_operator mode %d: %.1f us" % (mode, k * 1e3), flush=True)
        assert w > 0 and 0 < kmin <= k <= kmax
    with pytest.raises(ta.TransferError):
        ctx.bench_operator(1, 3, precision=5)
