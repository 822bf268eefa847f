"""GPU parity: the HIP operator, BLAS-1, reductions and PCG through the C ABI
against the FP64 oracle on the same inputs.

Bar (FP64): every operator output within 1e-10 relative to max|reference|
(observed ~1e-15); copies exact; PCG iterates (alpha, beta, ||r||) within
1e-9 relative over the first iterations, solutions within the convergence
tolerance's reach."""
import numpy as np
import pytest

import oracle
import transfer_amd as ta
from conftest import ang_dict, ref_angular_fixture

pytestmark = pytest.mark.gpu
TOL = 1e-10


def cases():
    return {
        "slab12-K1-iso": lambda: ta.problems.slab(),
        "ball64-K3": lambda: ta.problems.ball(64),
        "ball16-K5": lambda: ta.problems.ball(16, maxk=5),
        "ball12-K2-hg": lambda: ta.problems.ball(12, maxk=2, kind="hg"),
        "cube8-K3-iso": lambda: ta.problems.cube(8),
        # the reference's limit slm <= ASLM_MAX = 128 (common.cuh:7): K = 6, 7
        "ball10-K6": lambda: ta.problems.ball(10, maxk=6),
        "ball10-K7-hg": lambda: ta.problems.ball(10, maxk=7, kind="hg"),
    }


_cache = {}


def setup(name):
    if name not in _cache:
        print("[gpu-parity] building", name, flush=True)
        P = cases()[name]()
        ctx = ta.HipContext(P.mesh, P.ang)
        op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
        _cache[name] = (P, ctx, op)
    return _cache[name]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rvec(op, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((op.nP, op.aslm))
    f[:, :op.slm] = rng.standard_normal((op.nP, op.slm))
    return f.reshape(-1)


@pytest.mark.parametrize("name", list(cases()))
def test_lhs_lhsdiag_rhs_parity(name):
    P, ctx, op = setup(name)
    assert ctx.N() == op.N()
    f = rvec(op, 1)
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    want = op.lhs(f)
    default = ctx.schedule()
    for variant in (1, 2, 3):
        for J in sorted({default["J"], 1 if op.slm <= 64 else default["J"]}):
            if not try_schedule(ctx, 0, J=J, WPB=4, variant=variant):
                # the LDS-staged schedules hold the angular tables; variant 1 always fits
                print("[gpu-parity]", name, "J=%d variant=%d not available" % (J, variant), flush=True)
                assert variant != 1
                continue
            ctx.computeLhs(d_f, d_r)
            r = ctx.copyToHost(d_r)
            print("[gpu-parity]", name, "J=%d variant=%d lhs rel err %.2e" % (J, variant, rel(r, want)), flush=True)
            assert rel(r, want) <= TOL
    for wpb in (1, 4, 8):                       # matrix-core schedule, FP64 MFMA
        if not try_schedule(ctx, 0, J=0, WPB=wpb, variant=4):
            continue
        ctx.computeLhs(d_f, d_r)
        e = rel(ctx.copyToHost(d_r), want)
        print("[gpu-parity]", name, "variant=4 WPB=%d lhs rel err %.2e" % (wpb, e), flush=True)
        assert e <= TOL
    ctx.set_schedule(0, J=default["J"], WPB=default["WPB"], variant=default["variant"])
    ctx.computeLhs(d_f, d_r)
    r = ctx.copyToHost(d_r)
    pad = r.reshape(op.nP, op.aslm)[:, op.slm:]
    assert np.all(pad == 0)
    ctx.computeLhsDiag(d_f, d_r)
    assert rel(ctx.copyToHost(d_r), op.lhs_diag(f)) <= TOL
    ctx.computeRhs(d_r)
    assert rel(ctx.copyToHost(d_r), op.rhs()) <= 1e-14
    for v in (d_f, d_r):
        ctx.deviceFree(v)


def try_schedule(ctx, precision, **kw):
    """Forces a schedule; False where the mesh or harmonic count rules it out
    (the matrix-core schedule needs ceil(slm/4) in {2, 4, 7} and 15 or 16
    stencil slots)."""
    try:
        ctx.set_schedule(precision, **kw)
        return True
    except ta.TransferError:
        return False


@pytest.mark.parametrize("name", list(cases()))
def test_fp32_operator_every_schedule(name):
    """The FP32 operator (inner operator of the mixed solver) against the FP64
    oracle: bar 1e-5 relative to max|A f| (FP32 rounding of coefficients,
    inputs and sums; observed ~1e-7)."""
    P, ctx, op = setup(name)
    f = rvec(op, 5)
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    want = op.lhs(f)
    # largest instantiated J (harmonics per lane) of the VALU schedules
    J0 = max(j for j in (1, 2, 3, 4, 5, 6, 7) if op.slm % j == 0 and op.slm // j <= 64)
    for variant in (1, 2, 3):
        if not try_schedule(ctx, 1, J=J0, WPB=4, variant=variant):
            print("[gpu-parity] %s fp32 J=%d variant=%d not available" % (name, J0, variant), flush=True)
            assert variant != 1
            continue
        ctx.computeLhsPrec(1, d_f, d_r)
        e = rel(ctx.copyToHost(d_r), want)
        print("[gpu-parity] %s fp32 J=%d variant=%d rel err %.2e" % (name, J0, variant, e), flush=True)
        assert e <= 1e-5
    ctx.set_angular_precision(0)                      # FP32 MFMA projection
    try:
        if try_schedule(ctx, 1, J=0, WPB=4, variant=4):
            ctx.computeLhsPrec(1, d_f, d_r)
            e = rel(ctx.copyToHost(d_r), want)
            print("[gpu-parity] %s fp32 variant=4 (fp32 mfma) rel err %.2e" % (name, e), flush=True)
            assert e <= 1e-5
    finally:
        ctx.set_angular_precision(3)
    ctx.deviceFree(d_f)
    ctx.deviceFree(d_r)


@pytest.mark.parametrize("name", list(cases()))
def test_bf16_angular_projection(name):
    """The FP32 operator with the angular projection on the BF16 matrix cores
    (FP32 accumulation) against the FP64 oracle.  Plain BF16 rounds the
    gathered values and the couplings to 8 significant bits: bar 2e-2
    relative to max|A f|.  The hi/lo split (three BF16 products per term, the
    lo*lo term dropped) leaves ~2^-16 per term: bar 2e-4."""
    P, ctx, op = setup(name)
    f = rvec(op, 6)
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    want = op.lhs(f)
    try:
        for prec, bar in ((1, 2e-2), (3, 2e-4)):
            ctx.set_angular_precision(prec)
            if not try_schedule(ctx, 1, J=0, WPB=4, variant=4):
                # no 16x16 matrix-core tile of this order: 33..80 harmonics take
                # the K-blocked schedule (variant 5, the same bars), more take
                # the exact FP32 VALU schedule (bar 1e-5)
                blocked = try_schedule(ctx, 1, J=0, WPB=4, variant=5)
                if not blocked:
                    ctx.autotune(1)
                ctx.computeLhsPrec(1, d_f, d_r)
                e = rel(ctx.copyToHost(d_r), want)
                print("[gpu-parity] %s angular prec=%d via %s rel err %.2e" % (name, prec, ctx.kernel_name(1), e),
                      flush=True)
                assert e <= (bar if blocked else 1e-5)
                continue
            for wpb in (1, 4, 8):
                ctx.set_schedule(1, J=0, WPB=wpb, variant=4)
                ctx.computeLhsPrec(1, d_f, d_r)
                e = rel(ctx.copyToHost(d_r), want)
                print("[gpu-parity] %s bf16 prec=%d WPB=%d rel err %.2e" % (name, prec, wpb, e), flush=True)
                assert e <= bar
    finally:
        ctx.set_angular_precision(3)
        ctx.deviceFree(d_f)
        ctx.deviceFree(d_r)


@pytest.mark.parametrize("name", ["ball16-K5", "ball10-K5-hg", "ball12-K4"])
def test_bf16_blocked_projection(name):
    """The FP32 operator for 33..80 harmonics with the BF16 matrix-core
    projection in blocks of at most 16 harmonics (variant 5) against the
    FP64 oracle.  Plain BF16 (gathered values and couplings rounded to 8
    significant bits: bar 2e-2 relative to max|A f|, observed 3e-3) or the
    hi/lo split of both (bar 2e-4, observed 6e-6); every workgroup size,
    under the static class layout (straight-line contraction, k_apply5s: the
    default where it applies) and the masked block layout (k_apply5).
    Angular precision 4: FP32 blocks and products on the FP32 matrix cores
    under the static layout (bar 1e-6, observed 2e-7 to 3e-7), the hi/lo
    split under the masked one."""
    if name in cases():
        P, ctx, op = setup(name)
    else:
        P = ta.problems.ball(10, maxk=5, kind="hg") if name == "ball10-K5-hg" else ta.problems.ball(12, maxk=4)
        ctx = ta.HipContext(P.mesh, P.ang)
        op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    f = rvec(op, 7)
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    want = op.lhs(f)
    try:
        for layout in (0, 1):
            for prec, bar in ((1, 2e-2), (3, 2e-4), (4, 1e-6 if layout == 0 else 2e-4)):
                ctx.set_angular_precision(prec)
                for wpb in (2, 4, 8):
                    ctx.set_schedule(1, J=layout, WPB=wpb, variant=5)
                    kname = ctx.kernel_name(1)
                    # the classes' static layout covers K = 4, 5 (seven or ten channels)
                    assert kname.startswith("k_apply5s<" if layout == 0 else "k_apply5<"), kname
                    if layout == 0:
                        assert kname.endswith(", %d, 0>" % (4 if prec == 4 else 3 if prec == 3 else 1)), kname
                    ctx.computeLhsPrec(1, d_f, d_r)
                    e = rel(ctx.copyToHost(d_r), want)
                    print("[gpu-parity] %s bf16-blocked prec=%d WPB=%d rel err %.2e (%s)" % (
                        name, prec, wpb, e, kname), flush=True)
                    assert e <= bar
    finally:
        ctx.set_schedule(1, J=0, WPB=8, variant=5)
        ctx.set_angular_precision(3)
        ctx.deviceFree(d_f)
        ctx.deviceFree(d_r)


@pytest.mark.parametrize("name", ["ball16-K5", "ball10-K5-hg"])
def test_fp64_blocked_operator_and_solves(name):
    """The FP64 operator for 33..80 harmonics on the FP64 matrix cores in
    parity-class blocks (variant 5 in FP64; seven channels in one group,
    ten in two) against the oracle at the FP64 bar, the angular-diagonal
    operator beside it, and FP64 PCG solves through its fused epilogues
    (p.Ap with Jacobi, the Chebyshev steps) against the VALU schedule's: the
    oracle re-checks every solution's residual and the iteration counts
    agree to rounding (within 2%)."""
    if name in cases():
        P, ctx, op = setup(name)
    else:
        P = ta.problems.ball(10, maxk=5, kind="hg")
        ctx = ta.HipContext(P.mesh, P.ang)
        op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    default = ctx.schedule()
    f = rvec(op, 11)
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    b = op.rhs()
    try:
        assert try_schedule(ctx, 0, J=0, WPB=8, variant=5), "FP64 K-blocked schedule expected here"
        ctx.computeLhs(d_f, d_r)
        e = rel(ctx.copyToHost(d_r), op.lhs(f))
        print("[gpu-parity] %s fp64-blocked %s lhs rel err %.2e" % (name, ctx.kernel_name(0), e), flush=True)
        assert e <= TOL
        ctx.computeLhsDiag(d_f, d_r)
        assert rel(ctx.copyToHost(d_r), op.lhs_diag(f)) <= TOL
        iters = {}
        for use_prec in (0, 2, 4):
            for variant in (5, default["variant"]):
                if variant == 5:
                    ctx.set_schedule(0, J=0, WPB=8, variant=5)
                else:
                    ctx.set_schedule(0, J=default["J"], WPB=default["WPB"], variant=default["variant"])
                d_x = ctx.allocVector()
                st, _ = ctx.solve(d_x, solver=ta.TR_SOLVER_PCG_FP64, use_prec=use_prec, maxit=5000, tol_abs=0.0,
                                  tol_rel=1e-10)
                res = op.norm(b - op.lhs(ctx.copyToHost(d_x))) / op.norm(b)
                ctx.deviceFree(d_x)
                print("[gpu-parity] %s fp64 PCG use_prec=%d variant=%d: %d iterations, true rel %.2e, oracle %.2e" % (
                    name, use_prec, variant, st.iters, st.true_rel, res), flush=True)
                assert st.status == 0 and res <= 1.5e-10
                iters[(use_prec, variant)] = st.iters
            a, c = iters[(use_prec, 5)], iters[(use_prec, default["variant"])]
            assert abs(a - c) <= max(2, 0.02 * c)
    finally:
        ctx.set_schedule(0, J=default["J"], WPB=default["WPB"], variant=default["variant"])
        ctx.deviceFree(d_f)
        ctx.deviceFree(d_r)


def test_copies_blas_reductions():
    P, ctx, op = setup("slab12-K1-iso")
    x, y = rvec(op, 2), rvec(op, 3)
    dx, dy = ctx.vector_from(x), ctx.vector_from(y)
    assert np.array_equal(ctx.copyToHost(dx), x)           # permuting copies round-trip exactly
    small = ctx.getHostSmall(dx)
    assert np.array_equal(small, x.reshape(op.nP, op.aslm)[:, :op.slm].reshape(-1))
    assert abs(ctx.dot(dx, dy) - op.dot(x, y)) <= 1e-12 * abs(op.dot(np.abs(x), np.abs(y)))
    assert abs(ctx.norm(dx) - op.norm(x)) <= 1e-13 * op.norm(x)
    ctx.scale(dx, 0.5)
    ctx.addProd(dx, dy, -2.0)
    ctx.mulAdd(dx, 3.0, dy)
    ctx.mulAddProd(dx, -1.5, dy, 0.25)
    want = -1.5 * (3.0 * (0.5 * x - 2.0 * y) + y) + 0.25 * y
    np.testing.assert_allclose(ctx.copyToHost(dx), want, rtol=1e-15, atol=1e-15)
    raw = ctx.deviceAlloc(64)
    host = np.arange(8, dtype=np.float64)
    ctx.copyToDev(raw, host)
    assert np.array_equal(ctx.copyToHost(raw, 64), host)  # raw buffers are copied byte for byte
    for v in (dx, dy, raw):
        ctx.deviceFree(v)


@pytest.mark.parametrize("use_prec", [False, True])
def test_pcg_parity_reference_algorithm(use_prec):
    P, ctx, op = setup("slab12-K1-iso")
    kw = dict(use_prec=use_prec, maxit=400, tol_abs=0.0, tol_rel=1e-10, inner_tol=1e-10, record=6)
    f_or, st_or, h_or = op.pcg(**kw)
    d_f = ctx.allocVector()
    st, h = ctx.solve(d_f, solver=ta.TR_SOLVER_REFERENCE, **kw)
    f = ctx.copyToHost(d_f)
    print("[gpu-parity] pcg prec=%d iters gpu %d oracle %d preck %d/%d true_rel %.2e" %
          (use_prec, st.iters, st_or.iters, st.preck, st_or.preck, st.true_rel), flush=True)
    np.testing.assert_allclose(h, h_or, rtol=1e-9)
    assert st.status == 0 and st.true_rel <= 2e-10
    assert abs(st.iters - st_or.iters) <= 2
    assert rel(f, f_or) <= 1e-8
    ctx.deviceFree(d_f)


@pytest.mark.parametrize("K", [3, 7])
def test_reference_angular_data_through_the_view(K):
    """The context accepts the reference's own AngularData arrays (the
    committed fixture written from its sources) and matches the oracle fed
    the same arrays."""
    ref = ref_angular_fixture(K)
    P = ta.problems.ball(12)
    av = ta.Angular.view_of(ref["omega"], ref["omega_pos"], ref["Ox"], ref["Oy"], ref["Oz"])
    ctx = ta.HipContext(P.mesh, None, ang_view=av)
    op = oracle.OracleProblem(P.mesh.arrays(), ref)
    f = rvec(op, 4)
    d_f, d_r = ctx.vector_from(f), ctx.allocVector()
    ctx.computeLhs(d_f, d_r)
    assert rel(ctx.copyToHost(d_r), op.lhs(f)) <= TOL
    ctx.close()


@pytest.mark.parametrize("solver", [ta.TR_SOLVER_REFERENCE, ta.TR_SOLVER_PCG_FP64])
def test_pcg_iterations_match_oracle_ball32(solver):
    """main.cpp:125-162's PCG (no invprec) on a K = 3 Planck ball of the
    64^3 ball's construction at 32^3: the host-driven reference algorithm
    and the device-resident PCG both take the oracle's iteration count to a
    1e-8 relative residual (within 2: the sums run in another order) and
    reach its solution to 1e-7."""
    P = ta.problems.ball(32)
    ctx = ta.HipContext(P.mesh, P.ang)
    op = oracle.OracleProblem(P.mesh.arrays(), ang_dict(P.ang))
    kw = dict(maxit=5000, tol_abs=0.0, tol_rel=1e-8)
    f_or, st_or, _ = op.pcg(**kw)
    d_f = ctx.allocVector()
    st, _ = ctx.solve(d_f, solver=solver, **kw)
    f = ctx.copyToHost(d_f)
    print("[gpu-parity] ball32 pcg solver=%d iters gpu %d oracle %d, solution rel diff %.2e" % (
        solver, st.iters, st_or.iters, rel(f, f_or)), flush=True)
    assert st.status == 0 and abs(st.iters - st_or.iters) <= 2
    assert rel(f, f_or) <= 1e-7
    ctx.close()


def test_padding_lanes_of_inputs_are_ignored():
    """Vectors cross the boundary in the reference layout [v][aslm]; the
    library keeps [v][slm] inside, so padding lanes of an input are dropped
    and every output's padding lanes are zero.  (The reference's vertex
    kernel would carry nonzero input padding into the kappa V / 20 mass of
    the output's padding lanes, kernels.cu:244-249, 286; its callers never
    make such vectors: allocations are zero-filled and the BLAS-1 kernels
    map zero padding to zero.)  The real lanes equal the oracle's operator
    applied to the input with its padding cleared."""
    P, ctx, op = setup("ball12-K2-hg")
    f = rvec(op, 9).reshape(op.nP, op.aslm)
    f[:, op.slm:] = 7.0
    d_f, d_r = ctx.vector_from(f.ravel()), ctx.allocVector()
    ctx.computeLhs(d_f, d_r)
    r = ctx.copyToHost(d_r).reshape(op.nP, op.aslm)
    g = f.copy()
    g[:, op.slm:] = 0.0
    want = op.lhs(g.ravel()).reshape(op.nP, op.aslm)
    assert np.all(r[:, op.slm:] == 0.0)
    assert rel(r[:, :op.slm], want[:, :op.slm]) <= TOL
    ctx.deviceFree(d_f)
    ctx.deviceFree(d_r)
