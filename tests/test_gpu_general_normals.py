"""GPU: vacuum faces off the axes (TR_CTX_GENERAL_NORMALS; the reference's
axis rule, kernels.cu:359-361, is exact only along the axes, Issue #15).

The pin is rotation covariance.  A mesh B that is mesh A rotated by R poses
the same problem in rotated directions, so the harmonic coefficients at
every vertex transform by the rotation T of the even harmonics up to 2K,
T[lm][l'm'] = <Y_lm(Omega) Y_l'm'(R^T Omega)> (exact: harmonics of degree
<= 2K are closed under rotation, and the rule below integrates degree 4K
exactly).  With general normals A_B (T f) = T (A_A f) at 1e-12 relative to
max|A f|, for a generic rotation (every boundary face off the axes) and a
rotation about z (the z faces stay on an axis): both boundary paths at once.
Solutions then agree in U = 4 pi f_00 (rotation invariant) to the solver
tolerance, and without the flag (the reference's rule) they do not."""
import numpy as np
import pytest

import transfer_amd as ta
from angular_np import harmonics, sphere_rule

pytestmark = pytest.mark.gpu
K = 2


def rotation(kind):
    if kind == "z30":
        c, s = np.cos(np.pi / 6), np.sin(np.pi / 6)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    a, b, g = 0.4, 1.1, -0.7
    Rz = lambda t: np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    Ry = lambda t: np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    return Rz(a) @ Ry(b) @ Rz(g)


def sh_rotation(R, K=K):
    mu, ph, w = sphere_rule(2 * K + 8, 4 * K + 16)
    s = np.sqrt(1 - mu ** 2)
    d = np.stack([s * np.cos(ph), s * np.sin(ph), mu])
    dr = R.T @ d
    ls = list(range(0, 2 * K + 1, 2))
    Y = harmonics(ls, d[2], np.arctan2(d[1], d[0]))
    Yr = harmonics(ls, np.clip(dr[2], -1, 1), np.arctan2(dr[1], dr[0]))
    return (Y * w) @ Yr.T


@pytest.fixture(scope="module")
def meshes():
    A = ta.Mesh.box((4, 4, 4), (0, 0, 0), (1, 1, 1), jitter=0.2, seed=5)
    coords, tets, cen, vol = A.geometry()
    kappa = 1.0 + 2.0 * ta.rng_uniform(3, 1, len(tets))
    Ip = np.ones(len(tets))
    A.set_fields(kappa=kappa, Ip=Ip)
    return A, coords, tets, kappa, Ip


# K = 2 (slm 15): two faces per wavefront, O_z in LDS; K = 4, 5 (slm 45,
# 66): one face per wavefront; K = 7 (slm 120): a workgroup of 128 lanes per
# face with O_z read from global memory
@pytest.mark.parametrize("kind,K", [("generic", 2), ("z30", 2), ("generic", 4), ("generic", 5), ("z30", 5),
                                    ("generic", 7)])
def test_operator_is_rotation_covariant(meshes, kind, K):
    A, coords, tets, kappa, Ip = meshes
    R = rotation(kind)
    B = ta.Mesh.from_arrays(coords @ R.T, tets, kappa, Ip)
    ang = ta.Angular(K)
    print("[gpu-normals] covariance %s K=%d" % (kind, K), flush=True)
    ctxA, ctxB = ta.HipContext(A, ang), ta.HipContext(B, ang, general_normals=True)
    T = sh_rotation(R, K)
    assert np.abs(T @ T.T - np.eye(len(T))).max() < 1e-12
    nP, slm, aslm = A.nP, ctxA.slm, ctxA.aslm
    rng = np.random.default_rng(1)
    fA = np.zeros((nP, aslm))
    fA[:, :slm] = rng.standard_normal((nP, slm))
    fB = np.zeros((nP, aslm))
    fB[:, :slm] = fA[:, :slm] @ T.T
    dA, dB, rA, rB = ctxA.vector_from(fA.ravel()), ctxB.vector_from(fB.ravel()), ctxA.allocVector(), ctxB.allocVector()
    ctxA.computeLhs(dA, rA)
    ctxB.computeLhs(dB, rB)
    want = ctxA.copyToHost(rA).reshape(nP, aslm)[:, :slm] @ T.T
    got = ctxB.copyToHost(rB).reshape(nP, aslm)[:, :slm]
    err = np.abs(got - want).max() / np.abs(want).max()
    ctxB.set_angular_precision(0)                  # exact FP32 couplings (no BF16 rounding)
    ctxB.computeLhsPrec(1, dB, rB)
    err32 = np.abs(ctxB.copyToHost(rB).reshape(nP, aslm)[:, :slm] - want).max() / np.abs(want).max()
    # the Diag operator with the general faces (gOd = diag(R O_z R^T)): its
    # diagonal is A's, probed with unit vectors at boundary vertices
    bv = np.unique(B.arrays()["bnd"]["p"][:, :3])[:3]
    for v in bv:
        for lm in sorted({0, slm // 2, slm - 1}):
            e = np.zeros((nP, aslm))
            e[v, lm] = 1.0
            de = ctxB.vector_from(e.ravel())
            ctxB.computeLhs(de, rB)
            a_ii = ctxB.copyToHost(rB).reshape(nP, aslm)[v, lm]
            ctxB.computeLhsDiag(de, rB)
            d_ii = ctxB.copyToHost(rB).reshape(nP, aslm)[v, lm]
            assert abs(d_ii - a_ii) <= 1e-12 * abs(a_ii), (v, lm, d_ii, a_ii)
            ctxB.deviceFree(de)
    # the reference's axis rule on the rotated mesh
    ctx0 = ta.HipContext(B, ang)
    d0, r0 = ctx0.vector_from(fB.ravel()), ctx0.allocVector()
    ctx0.computeLhs(d0, r0)
    err0 = np.abs(ctx0.copyToHost(r0).reshape(nP, aslm)[:, :slm] - want).max() / np.abs(want).max()
    print("[gpu-normals] %s: A_B T f vs T A_A f rel err %.2e (FP32 operator %.2e); axis rule %.2e" % (
        kind, err, err32, err0), flush=True)
    assert err <= 1e-12 and err32 <= 1e-5 and err0 > 1e-3


def test_solutions_agree_in_scalar_radiation(meshes):
    A, coords, tets, kappa, Ip = meshes
    R = rotation("generic")
    B = ta.Mesh.from_arrays(coords @ R.T, tets, kappa, Ip)
    ang = ta.Angular(K)
    out = {}
    # Chebyshev steps fused into the matrix-core ("B-*-cheb") and the
    # buffer-load VALU ("-v3") operators, the faces' share added after them
    for name, mesh, gen, solver, prec in (("A", A, False, ta.TR_SOLVER_PCG_FP64, 2),
                                         ("B", B, True, ta.TR_SOLVER_PCG_FP64, 2),
                                         ("B-fp64-cheb", B, True, ta.TR_SOLVER_PCG_FP64, 4),
                                         ("B-mixed-cheb", B, True, ta.TR_SOLVER_MIXED, 4),
                                         ("B-mixed-cheb-v3", B, True, ta.TR_SOLVER_MIXED, 4),
                                         ("B-fp64-multilevel", B, True, ta.TR_SOLVER_PCG_FP64, 3),
                                         ("B-axis-rule", B, False, ta.TR_SOLVER_PCG_FP64, 2)):
        ctx = ta.HipContext(mesh, ang, general_normals=gen)
        if name.endswith("multilevel"):
            ctx.set_multilevel(16, 8, 2)  # coarse levels even on this small mesh
        if name.endswith("-v3"):
            for p in (0, 1):
                ctx.set_schedule(p, J=3, WPB=4, variant=3)  # slm 15 at K = 2
        d = ctx.allocVector()
        st, _ = ctx.solve(d, solver=solver, use_prec=prec, maxit=5000 if solver != ta.TR_SOLVER_MIXED else 30,
                          tol_abs=0.0, tol_rel=1e-11 if solver != ta.TR_SOLVER_MIXED else 1e-9, inner_maxit=5000,
                          inner_tol=1e-4)
        assert st.status == 0
        out[name] = ctx.moments(d)[0]
        print("[gpu-normals] %s: %d iterations (%d inner), true rel %.2e" % (name, st.iters, st.preck, st.true_rel),
              flush=True)
    ref = np.abs(out["A"]).max()
    e = {k: np.abs(out[k] - out["A"]).max() / ref for k in out if k != "A"}
    print("[gpu-normals] U against the unrotated solve: %s" % e, flush=True)
    assert e["B"] <= 1e-8 and e["B-fp64-cheb"] <= 1e-8 and e["B-axis-rule"] > 1e-3
    assert e["B-mixed-cheb"] <= 1e-6 and e["B-mixed-cheb-v3"] <= 1e-6 and e["B-fp64-multilevel"] <= 1e-8


def test_batch_with_general_normals(meshes):
    """A batch of the mesh, its generic rotation and its z rotation with
    general normals: the union operator restricted to each problem equals
    that problem's own context (1e-12), and the batched Chebyshev solve gives
    each problem's U of the solve alone (1e-8), which is also the unrotated
    one's."""
    A, coords, tets, kappa, Ip = meshes
    ms = [A] + [ta.Mesh.from_arrays(coords @ rotation(k).T, tets, kappa, Ip) for k in ("generic", "z30")]
    ang = ta.Angular(K)
    bat = ta.Batch(ms, ang, [0.0] * len(ms), general_normals=True)
    rng = np.random.default_rng(9)
    f = np.zeros((bat.nP, bat.aslm))
    f[:, :bat.slm] = rng.standard_normal((bat.nP, bat.slm))
    d_f, d_r = bat.vector_from(f.ravel()), bat.allocVector()
    bat.computeLhs(d_f, d_r)
    r = bat.copyToHost(d_r).reshape(bat.nP, bat.aslm)
    d_x = bat.allocVector()
    st, per = bat.solve_batch(d_x, solver=ta.TR_SOLVER_PCG_FP64, use_prec=4, maxit=5000, tol_rel=1e-11)
    assert st.status == 0
    x = bat.copyToHost(d_x)
    Ub = bat.moments(d_x)[0]
    for p, m in enumerate(ms):
        lo, hi = int(bat.voff[p]), int(bat.voff[p + 1])
        one = ta.HipContext(m, ang, general_normals=True)
        o_f, o_r = one.vector_from(f[lo:hi].ravel()), one.allocVector()
        one.computeLhs(o_f, o_r)
        want = one.copyToHost(o_r).reshape(hi - lo, one.aslm)
        e_op = np.abs(r[lo:hi] - want).max() / np.abs(want).max()
        d = one.allocVector()
        s1, _ = one.solve(d, solver=ta.TR_SOLVER_PCG_FP64, use_prec=2, maxit=5000, tol_abs=0.0, tol_rel=1e-11)
        U1 = one.moments(d)[0]
        e_u = np.abs(Ub[lo:hi] - U1).max() / np.abs(U1).max()
        print("[gpu-normals] batch problem %d: operator %.2e, U vs alone %.2e (%d iterations)" % (
            p, e_op, e_u, per[p]["iters"]), flush=True)
        assert e_op <= 1e-12 and e_u <= 1e-8
        one.close()
    assert np.isfinite(x).all()
