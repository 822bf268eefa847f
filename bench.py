#!/usr/bin/env python3
"""Benchmark of the radiative-transfer operator and its preconditioned solve
on one MI355X.

Default workload (the high-order scattering configuration, the largest that
fits one GPU): the 96^3-grid staircase ball (cubes of a 96^3 grid on
[-1,1]^3 with centroid r <= 1: 2,780,400 tets, 485,385 vertices), K = 5
(66 harmonics, the 194-point Lebedev order), Henyey-Greenstein scattering
g = 0.6 truncated after P_7 with albedo 0.9, kappa = 1 + 9 U[0,1) per tet,
I_p = 1.

A step is one application of the FP32 operator of the mixed solver (FP32
gather, BF16 matrix-core angular projection with FP32 accumulation) to a
device-resident vector.  value = DOF*ordinate/s = nP * slm * steps / wall
seconds of the timed loop (device-synchronised on both sides).  Also
reported: the roofline of that kernel (HIP events on its stream), the FP64
operator and the exact FP32 operator beside it, each operator's error
against FP64, the time to a 1e-8 relative FP64 residual of the plain FP64
GPU solver (the reference's PCG, main.cpp:125-162, on the device), of the
reference's invprec-preconditioned PCG (main.cpp:223-248), of the mixed
solvers and of the learned guess in front of the mixed solver; the 64^3
Planck ball as a secondary workload; the reference's CPU configuration
(the 12x12x48 slab) solved on the GPU; and the oracle timed on the host
cores.  --c2 runs the 64^3 Planck ball as the main workload, --batch N the
batching workload, --sphere N the curved sphere, each as its own JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-1e-8 residual & DOF·ordinate/s, learned+mixed vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0
CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT = 6, 300.0, 2


def operator_bytes(nT, nP, nF, slm, p):
    """Algorithmic bytes of one operator application (the tet-centric
    minimum): per tet 4 u32 indices + 1/(kappa V) + 3 stored s rows
    (16 + 10p), f read and r written once (2 nP slm p), per boundary face
    3 u32 indices + its area vector's axis weight (12 + p)."""
    return nT * (16 + 10 * p) + 2 * nP * slm * p + nF * (12 + p)


def source_digest():
    """sha256 (first 16 hex digits) of the operator kernels' sources: a
    counter summary measured on other sources does not describe the kernel
    that runs."""
    import hashlib
    h = hashlib.sha256()
    for name in ("kernels_mfma.hip", "kernels_mfma5s.hip", "mfma_common.h", "kernels.hip", "kernels.h"):
        with open(os.path.join(ROOT, "transfer_amd", "csrc", name), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel, workload, wpb=None):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, each scaled by the
    factor the same counter reads on a plain copy of known size).  Rows
    carry the source digest they were measured on and the workgroup size
    ("WPB n": the traffic depends on it); (None, reason) when no row covers
    this kernel at this workgroup size on these sources."""
    fn = os.path.join(ROOT, "profiles", "traffic_%s.txt" % workload)
    if not os.path.exists(fn):
        return None, "no counter summary for %s" % workload
    dig = source_digest()
    why = "kernel not in the summary"
    for line in open(fn):
        parts = [x.strip() for x in line.split("|")]
        if len(parts) >= 4 and parts[0] == kernel:
            if parts[1] != dig:
                why = "summary measured on other kernel sources"
                continue
            if wpb is not None and (len(parts) < 8 or parts[2] != "WPB %d" % wpb):
                why = "summary measured at another workgroup size"
                continue
            try:
                return float(parts[-1]), "profiles/traffic_%s.txt" % workload
            except ValueError:
                return None, "unreadable row"
    return None, why


def _oracle_problem(P, nthreads=0):
    import oracle
    a = P.ang
    return oracle.OracleProblem(P.mesh.arrays(), dict(omega=a.omega, omega_pos=a.omega_pos, Ox=a.Ox, Oy=a.Oy,
                                                      Oz=a.Oz, massfac=a.massfac, slm=a.slm), nthreads=nthreads)


def _time_lhs(op, seconds, min_apps=1):
    f = op.rhs()
    n, t0 = 0, time.perf_counter()
    while n < min_apps or time.perf_counter() - t0 < seconds:
        op.lhs(f)
        n += 1
    return n, time.perf_counter() - t0


def cpu_baseline(P, small, c1, seconds=8.0):
    """The oracle (kernels.cu and main.cpp's PCG restated in C, OpenMP over
    vertices) on the host cores, bounded samples of the workloads: the
    operator on the full mesh with every thread and on a smaller mesh of the
    same construction with one thread; the PCG of main.cpp:125-162 with and
    without invprec (main.cpp:223-248) to a 1e-8 relative residual on the
    reference's own CPU configuration (the 12x12x48 slab)."""
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    op = _oracle_problem(P)
    op.lhs(op.rhs())
    n, dt = _time_lhs(op, seconds, 2)
    value = op.nP * op.slm * n / dt
    del op
    op1 = _oracle_problem(small, nthreads=1)
    n1, dt1 = _time_lhs(op1, seconds / 2, 1)
    one = op1.nP * op1.slm * n1 / dt1
    pcg = {}
    for threads in (nthreads, 1):
        opc = _oracle_problem(c1, nthreads=0 if threads == nthreads else 1)
        for up in ((False, True) if threads == nthreads else (False,)):
            t0 = time.perf_counter()
            _, st, _ = opc.pcg(use_prec=up, maxit=5000, tol_abs=0.0, tol_rel=1e-8, inner_maxit=1000, inner_tol=1e-8)
            pcg["%s_threads%d" % ("invprec" if up else "plain", threads)] = dict(
                seconds=time.perf_counter() - t0, iters=st.iters, inner_iters=st.preck)
    out = dict(value=value, unit="DOF*ordinate/s", cores=nthreads, kind="port", nproc=os.cpu_count(),
               one_thread=dict(value=one, sample="%d applications on %s (%d vertices), 1 thread, %.1f s"
                               % (n1, small.name, op1.nP, dt1)),
               sample="%d applications of the FP64 oracle operator on the full %s mesh with %d threads, %.1f s"
                      % (n, P.name, nthreads, dt))
    out["pcg_to_1e-8"] = dict(workload="%s (the reference's CPU configuration: %d vertices, K = %d)"
                              % (c1.name, c1.mesh.nP, c1.ang.maxk), **pcg)
    return out


def bench_batch(count, steps, warmup, solve=True, copy_calibration=False):
    """The batching workload: `count` independent slabs (n ~ U{8..32}),
    log-normal opacity, per-problem albedo, K = 2, solved together to a 1e-8
    relative FP64 residual per problem."""
    import transfer_amd as ta
    t0 = time.perf_counter()
    def progress(p):
        if (p + 1) % 32 == 0 or p + 1 == count:
            print("[bench-batch] %d/%d meshes, %.1f s" % (p + 1, count, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    B = ta.problems.batch(count, progress=progress)
    t1 = time.perf_counter()
    ctx = ta.Batch(B.meshes, B.ang, B.albedo)
    t2 = time.perf_counter()
    nP, slm = ctx.nP, ctx.slm
    nT = sum(m.nT for m in B.meshes)
    print("[bench-batch] %d problems, %d vertices, %d tets; meshes %.1f s, assembly+upload %.1f s"
          % (count, nP, nT, t1 - t0, t2 - t1), file=sys.stderr, flush=True)
    ctx.autotune(0)
    # FP32 products on the matrix cores: as fast as the BF16 split on the
    # union (2.49 against 2.50 ms) and exact to FP32, which the Chebyshev
    # steps need (mixed Chebyshev 4.84 s against 6.95 s)
    ctx.set_angular_precision(0)
    ctx.autotune(1)
    out = dict(metric="batched time-to-1e-8 residual & DOF*ordinate/s", unit="DOF*ordinate/s", n_gpus=1,
               steps=steps, warmup=warmup, higher_is_better=True, data="synthetic", dtype="f32+f64",
               config=dict(workload="batch%d: slabs n x n x 4n, n~U{8..32}, ln kappa~N(0,1), albedo~U[0,0.95], K=2"
                                    % count, nP=nP, nT=nT, slm=slm, problems=count))
    nF = sum(m.nF for m in B.meshes)
    for prec, key, p in ((0, "operator_fp64", 8), (1, "operator_fp32", 4)):
        w, k, _, _ = ctx.bench_operator(warmup, steps, precision=prec)
        b = operator_bytes(nT, nP, nF, slm, p)
        out[key] = dict(kernel=ctx.kernel_name(prec), kernel_ms=k, dof_ordinate_per_s=nP * slm * steps / w,
                        algorithmic_bytes=b, achieved_gbs=b / (k * 1e-3) / 1e9,
                        hbm_frac=b / (k * 1e-3) / 1e9 / HBM_PEAK_GBS)
        print("[bench-batch] %s %s %.3f ms (%.0f GB/s algorithmic, %.3f of HBM)" % (
            key, ctx.kernel_name(prec), k, out[key]["achieved_gbs"], out[key]["hbm_frac"]), file=sys.stderr, flush=True)
    out["value"] = out["operator_fp32"]["dof_ordinate_per_s"]
    out["ms_per_step"] = out["operator_fp32"]["kernel_ms"]
    if copy_calibration:
        cms = ctx.bench_copy(256 << 20, 10)
        print("[bench-batch] copy 256 MiB: %.3f ms" % cms, file=sys.stderr, flush=True)
    if not solve:
        print(json.dumps(out), flush=True)
        return
    solves = {}
    ctx.set_chebyshev(CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT)
    # the mixed solves' inner tolerances (residual replacement every 1e-2),
    # measured on this workload: Jacobi 5.18 s at 1e-2 (6.30 s at 3e-3),
    # Chebyshev 4.43 s at 1e-2 (4.97 s at 1e-3, 4.87 s at 1e-5), lambda_max
    # estimated beforehand
    for name, solver, prec, itol in (("batched_mixed_fp32_ir_jacobi", ta.TR_SOLVER_MIXED, 2, 1e-2),
                                     ("batched_mixed_fp32_ir_chebyshev", ta.TR_SOLVER_MIXED, 4, 1e-2),
                                     ("batched_pcg_fp64_jacobi", ta.TR_SOLVER_PCG_FP64, 2, 1e-4),
                                     ("batched_pcg_fp64_chebyshev", ta.TR_SOLVER_PCG_FP64, 4, 1e-4)):
        print("[bench-batch] solving: %s" % name, file=sys.stderr, flush=True)
        d_x = ctx.allocVector()
        st, per = ctx.solve_batch(d_x, solver=solver, use_prec=prec,
                                  maxit=20000 if solver != ta.TR_SOLVER_MIXED else 30, inner_tol=itol)
        solves[name] = dict(seconds=st.seconds, max_iters=st.iters, inner_iters_total=st.preck,
                            worst_true_rel=st.true_rel, converged=st.status == 0,
                            problems_converged=sum(p["status"] == 0 for p in per))
        print("[bench-batch] %s: %.3f s, worst true rel %.2e, %d/%d converged"
              % (name, st.seconds, st.true_rel, solves[name]["problems_converged"], count),
              file=sys.stderr, flush=True)
        ctx.deviceFree(d_x)
    out["time_to_1e-8"] = solves
    print(json.dumps(out), flush=True)


def bench_sphere(n, steps, warmup):
    """The curved-boundary workload: the two-region sphere of the soltype-1
    check (problems.sphere, every boundary face off the axes), the FP64
    operator with the reference's axis rule and with general normals, at
    K = 3 and 5."""
    import transfer_amd as ta
    out = dict(metric="FP64 operator time on a curved sphere mesh, axis rule vs general normals", unit="us",
               n_gpus=1, steps=steps, warmup=warmup, higher_is_better=False, data="synthetic", dtype="f64",
               config=dict(workload="sphere%d: %d^3 Kuhn cubes mapped onto the unit ball, core r < 0.3" % (n, n)))
    for K in (3, 5):
        P = ta.problems.sphere(n, maxk=K)
        for gen in (False, True):
            ctx = ta.HipContext(P.mesh, P.ang, general_normals=gen)
            ctx.autotune(0)
            w, k, _, _ = ctx.bench_operator(warmup, steps, precision=0)
            key = "K%d_%s" % (K, "general_normals" if gen else "axis_rule")
            out[key] = dict(operator_us_per_step=w / steps * 1e6, kernel=ctx.kernel_name(0), nP=ctx.nP,
                            nF=P.mesh.nF)
            print("[bench-sphere] %s: %.1f us per application (%s)" % (key, w / steps * 1e6, ctx.kernel_name(0)),
                  file=sys.stderr, flush=True)
            ctx.close()
    print(json.dumps(out), flush=True)


def log(tag, msg):
    print("[%s] %s" % (tag, msg), file=sys.stderr, flush=True)


def time_operator(ctx, P, prec, steps, warmup, p):
    """(wall s, dict) of `steps` timed applications of the operator of
    precision prec (0 FP64, 1 the FP32 operator of the mixed solver, 2 the
    FP64 operator on vectors in its own harmonic order: under the FP64
    K-blocked schedule, the kernel without the permutation of the input)."""
    nP, nT, slm = P.mesh.nP, P.mesh.nT, P.ang.slm
    w, k, kmin, kmax = ctx.bench_operator(warmup, steps, precision=prec)
    b = operator_bytes(nT, nP, P.mesh.nF, slm, p)
    q = 1 if prec == 1 else 0
    return w, dict(kernel=ctx.kernel_name(q), kernel_ms=k, kernel_ms_min=kmin, kernel_ms_max=kmax,
                   dof_ordinate_per_s=nP * slm * steps / w, algorithmic_bytes=b,
                   achieved_gbs=b / (k * 1e-3) / 1e9, schedule=ctx.schedule(q))


def run_solves(ctx, runs, tag, only=""):
    """Each run from f = 0 to a 1e-8 relative FP64 residual."""
    import transfer_amd as ta
    out = {}
    for name, solver, prec, kw in runs:
        if only and name != only:
            continue
        setup = kw.pop("setup", None)
        if setup:
            setup()
        ctx.set_residual_replacement(kw.pop("replacement", True))
        d = ctx.allocVector()
        p = dict(maxit=20000 if solver != ta.TR_SOLVER_MIXED else 30, tol_abs=0.0, tol_rel=1e-8, inner_maxit=20000,
                 inner_tol=1e-5)
        p.update(kw)
        st, _ = ctx.solve(d, solver=solver, use_prec=prec, **p)
        ctx.deviceFree(d)
        out[name] = dict(seconds=st.seconds, iters=st.iters, inner_iters=st.preck, true_rel=st.true_rel,
                         converged=st.status == 0, fallback=st.fallback)
        log(tag, "%s: %.4f s, %d iters (%d inner), true rel %.2e%s" % (
            name, st.seconds, st.iters, st.preck, st.true_rel, ", fallback %d" % st.fallback if st.fallback else ""))
    return out


def train_guess(train, tag, steps=800):
    """The learned guess's network (8 -> 64 -> 64 -> 1, BF16 weights)
    trained on solver-generated solutions of the problems `train`."""
    from transfer_amd import learned
    t0 = time.perf_counter()
    feat, targ, _ = learned.generate_single(train)
    g = learned.LearnedGuess(seed=3)
    hist = g.fit(feat, targ, steps=steps, batch=8192, lr=3e-3)
    info = dict(training_seconds=time.perf_counter() - t0, training_problems=len(train),
                training_samples=int(len(targ)), training_loss=float(hist[-10:].mean()))
    log(tag, "learned guess trained on %d problems (%d samples) in %.1f s, loss %.2e" % (
        len(train), len(targ), info["training_seconds"], info["training_loss"]))
    return g, info


def learned_mixed(ctx, P, g, solver_kw, tag):
    """The learned guess in front of the mixed solver.  The timed part
    builds the guess (features, network and guess vector on the device)
    and solves, after one untimed call (first-launch code loading and the
    context's feature scratch, as the operator's warm-up steps); the solver
    scales the guess to its energy-optimal multiple and falls back to a
    zero start when that is no better than 0."""
    import transfer_amd as ta
    ctx.deviceFree(g.write_guess(ctx, P.mesh, P.albedo, Ip=None))
    ctx.synchronize()
    t1 = time.perf_counter()
    d_f = g.write_guess(ctx, P.mesh, P.albedo, Ip=None)
    ctx.synchronize()
    guess_s = time.perf_counter() - t1
    kw = dict(maxit=30, tol_abs=0.0, tol_rel=1e-8, inner_maxit=20000, inner_tol=1e-5)
    kw.update(solver_kw)
    prec = kw.pop("use_prec")
    ctx.set_residual_replacement(kw.pop("replacement", True))
    st, _ = ctx.solve(d_f, solver=ta.TR_SOLVER_MIXED, use_prec=prec, x0=True, **kw)
    ctx.deviceFree(d_f)
    out = dict(seconds=guess_s + st.seconds, guess_seconds=guess_s, solve_seconds=st.seconds, iters=st.iters,
               inner_iters=st.preck, true_rel=st.true_rel, converged=st.status == 0, fallback=st.fallback)
    log(tag, "learned guess + mixed: %.4f s (guess %.4f s, solve %.4f s), %d inner, fallback %d, true rel %.2e" % (
        out["seconds"], guess_s, st.seconds, st.preck, st.fallback, st.true_rel))
    return out


def speedups(out, solves):
    plain = solves["pcg_fp64_device"]["seconds"]
    best_plain = min(v["seconds"] for k, v in solves.items() if k.startswith(("pcg_fp64", "reference"))
                     and v["converged"])
    mixed = min((v["seconds"], k) for k, v in solves.items() if "mixed" in k and v["converged"])
    out["speedup_vs_plain_fp64"] = plain / mixed[0]
    out["speedup_vs_best_plain_fp64"] = best_plain / mixed[0]
    out["best_mixed"] = mixed[1]
    if "learned_mixed" in solves and solves["learned_mixed"]["converged"]:
        out["speedup_learned_mixed_vs_plain_fp64"] = plain / solves["learned_mixed"]["seconds"]


def bench_c3(args):
    """The default line: the 96^3 HG ball, K = 5."""
    import numpy as np
    import transfer_amd as ta
    tag = "bench"
    t0 = time.perf_counter()
    P = ta.problems.ball(96, maxk=5, kind="hg")
    ctx = ta.HipContext(P.mesh, P.ang)
    nP, nT, slm = P.mesh.nP, P.mesh.nT, P.ang.slm
    log(tag, "%s nP=%d nT=%d nF=%d slm=%d, setup %.1f s" % (P.name, nP, nT, P.mesh.nF, slm,
                                                              time.perf_counter() - t0))
    ctx.autotune(0)
    ctx.set_angular_precision(0)
    ctx.autotune(1)
    exact32 = ctx.schedule(1)
    ctx.set_angular_precision(3)
    ctx.autotune(1)
    mfma32 = ctx.schedule(1)
    if args.copy_calibration:
        cms = ctx.bench_copy(256 << 20, 10)
        log(tag, "copy 256 MiB: %.3f ms, %.0f GB/s (read+write)" % (cms, 2 * (256 << 20) / cms / 1e6))
    steps, warmup = args.steps, args.warmup

    def use_exact():
        ctx.set_angular_precision(0)
        ctx.set_schedule(1, J=exact32["J"], WPB=exact32["WPB"], variant=exact32["variant"])

    def use_mfma():
        ctx.set_angular_precision(3)
        ctx.set_schedule(1, J=mfma32["J"], WPB=mfma32["WPB"], variant=mfma32["variant"])

    def use_f32mfma():
        ctx.set_angular_precision(4)
        ctx.set_schedule(1, J=mfma32["J"], WPB=mfma32["WPB"], variant=mfma32["variant"])

    wall, op32 = time_operator(ctx, P, 1, steps, warmup, 4)
    _, op64 = time_operator(ctx, P, 0, steps, warmup, 8)
    if op64["schedule"]["variant"] == 5:
        # the K-blocked FP64 kernel alone (its solves keep their vectors in
        # its order); op64 includes the input's permutation, as a caller of
        # computeLhs with internal-order vectors pays it
        _, k64 = time_operator(ctx, P, 2, steps, warmup, 8)
        op64["kernel_only"] = {k: k64[k] for k in ("kernel_ms", "kernel_ms_min", "kernel_ms_max", "achieved_gbs")}
        log(tag, "FP64 operator kernel alone %.1f us (%.0f GB/s algorithmic)" % (
            k64["kernel_ms"] * 1e3, k64["achieved_gbs"]))
    # the same FP32 operator as a solver iteration launches it: with the
    # fused p.Ap dot, and with a 1 GiB copy before each application (the
    # cache state that the iteration's vector passes leave behind)
    for mode, key in ((3, "fused_dot_ms"), (4, "fused_dot_after_copy_ms")):
        op32[key] = ctx.bench_operator(warmup, steps, precision=mode)[1]
    log(tag, "FP32/BF16 operator with the fused dot %.1f us, after a 1 GiB copy %.1f us" % (
        op32["fused_dot_ms"] * 1e3, op32["fused_dot_after_copy_ms"] * 1e3))
    use_exact()
    _, op32x = time_operator(ctx, P, 1, steps, warmup, 4)
    use_f32mfma()
    _, op32m = time_operator(ctx, P, 1, steps, warmup, 4)
    use_mfma()
    for k, v in (("FP32/BF16", op32), ("FP64", op64), ("exact FP32", op32x), ("FP32 matrix-core", op32m)):
        log(tag, "%s operator %s %.1f us (%.3g DOF*ordinate/s)" % (k, v["kernel"], v["kernel_ms"] * 1e3,
                                                                   v["dof_ordinate_per_s"]))
    # each FP32 operator's error against the FP64 one on a random vector
    f, a64, a32 = ctx.allocVector(), ctx.allocVector(), ctx.allocVector()
    ctx.copyToDev(f, np.random.default_rng(1).standard_normal(ctx.N()))
    ctx.computeLhs(f, a64)
    A64 = ctx.copyToHost(a64)
    for d, setup in ((op32x, use_exact), (op32m, use_f32mfma), (op32, use_mfma)):
        setup()
        ctx.computeLhsPrec(1, f, a32)
        d["rel_error_vs_fp64"] = float(np.linalg.norm(ctx.copyToHost(a32) - A64) / np.linalg.norm(A64))
    for v in (f, a64, a32):
        ctx.deviceFree(v)
    use_mfma()
    traffic, tsrc = measured_traffic(op32["kernel"], P.name, op32["schedule"]["WPB"])
    out = dict(metric=METRIC, value=op32["dof_ordinate_per_s"], unit="DOF*ordinate/s", n_gpus=1, steps=steps,
               warmup=warmup, ms_per_step=wall / steps * 1e3, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f32+bf16", data="synthetic",
               config=dict(workload="%s: ball of 96^3-grid cubes, K=5 (slm 66), HG g=0.6 P7, albedo 0.9, "
                                    "kappa=1+9U, I_p=1; step = one application of the mixed solver's FP32 "
                                    "operator (BF16 matrix-core projection)" % P.name,
                           nP=nP, nT=nT, nF=P.mesh.nF, slm=slm))
    out["roofline"] = dict(bound="hbm", achieved=op32["achieved_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=op32["achieved_gbs"] / HBM_PEAK_GBS, traffic=traffic, traffic_source=tsrc,
                           kernel=op32["kernel"], kernel_ms=op32["kernel_ms"],
                           algorithmic_bytes=op32["algorithmic_bytes"])
    out["operator_fp32_bf16"] = op32
    out["operator_fp64"] = op64
    out["operator_fp32_exact"] = op32x
    out["operator_fp32_matrix_core"] = op32m
    out["kernel_source_digest"] = source_digest()
    if not args.no_solve:
        # Chebyshev degree 6 over [lambda_max / 300, lambda_max], each step one
        # fused operator launch (the BF16 matrix-core operator's hi/lo split)
        ctx.set_chebyshev(CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT)
        # residual replacement (the default) with a replacement every 1e-2
        # of inner reduction; the restarted refinement (each step's inner CG
        # from zero to 1e-5) beside it
        mixed_kw = {"mixed_fp32_ir_jacobi": (use_exact, dict(use_prec=2, inner_tol=1e-2)),
                    "mixed_bf16mfma_ir_plain": (use_mfma, dict(use_prec=0, inner_tol=1e-2)),
                    "mixed_bf16mfma_ir_plain_restart": (use_mfma, dict(use_prec=0, inner_tol=1e-5,
                                                                       replacement=False)),
                    "mixed_bf16mfma_ir_jacobi": (use_mfma, dict(use_prec=2, inner_tol=1e-2)),
                    "mixed_bf16mfma_ir_chebyshev": (use_mfma, dict(use_prec=4, inner_tol=1e-4)),
                    "mixed_f32mfma_ir_plain": (use_f32mfma, dict(use_prec=0, inner_tol=1e-2))}
        runs = [("pcg_fp64_device", ta.TR_SOLVER_PCG_FP64, 0, {}),
                ("pcg_fp64_device_jacobi", ta.TR_SOLVER_PCG_FP64, 2, {}),
                ("pcg_fp64_device_chebyshev", ta.TR_SOLVER_PCG_FP64, 4, {})]
        for name, (setup, kw) in mixed_kw.items():
            kw = dict(kw)
            runs.append((name, ta.TR_SOLVER_MIXED, kw.pop("use_prec"), dict(setup=setup, **kw)))
        if args.invprec:
            runs.insert(3, ("reference_pcg_fp64_invprec_host", ta.TR_SOLVER_REFERENCE, 1,
                            dict(inner_tol=1e-8, inner_maxit=1000)))
        solves = run_solves(ctx, runs, tag, args.only)
        if not args.no_learned and not args.only:
            best = min((v["seconds"], k) for k, v in solves.items() if "mixed" in k and v["converged"])[1]
            setup, kw = mixed_kw[best]
            setup()
            train = [ta.problems.ball(n, maxk=5, kind="hg", seed=s) for n, s in
                     ((16, 11), (20, 12), (24, 13), (20, 14), (16, 15), (24, 16), (28, 17), (28, 18))]
            g, info = train_guess(train, tag)
            lm = learned_mixed(ctx, P, g, kw, tag)
            lm.update(info, inner_solver=best)
            solves["learned_mixed"] = lm
            use_mfma()
        out["time_to_1e-8"] = solves
        if not args.only:
            speedups(out, solves)
    if not args.only and not args.no_secondary:
        out["secondary_ball64"] = bench_c2_brief(min(steps, 50), min(warmup, 5))
        out["c1_reference_config_on_gpu"] = c1_on_gpu()
    if not args.no_cpu and not args.only:
        out["cpu_baseline"] = cpu_baseline(P, ta.problems.ball(32, maxk=5, kind="hg"), ta.problems.slab())
        log(tag, "cpu baseline %.3g DOF*ordinate/s (%d threads), 1 thread %.3g" % (
            out["cpu_baseline"]["value"], out["cpu_baseline"]["cores"], out["cpu_baseline"]["one_thread"]["value"]))
    print(json.dumps(out), flush=True)


def bench_c5(args):
    """The learned part's protocol on the 32^3 blob cube (196,608 tets,
    K = 3, kappa = 0.1 + 8 Gaussian blobs, albedo U[0, 0.99] per problem):
    the network trained on the solutions of args.c5_train solver-generated
    problems (seeds 0 ..), then, on args.c5_test held-out seeds, the mixed
    solver (Jacobi-preconditioned FP32 CG inside FP64 refinement) to a 1e-8
    FP64 residual from zero, from the learned guess, and from the guess with
    the learned coarse correction in the preconditioner (use_prec 5); the
    learned rows' time includes building the guess on the device."""
    import numpy as np
    import transfer_amd as ta
    from transfer_amd import learned
    tag = "bench-c5"
    ntrain, ntest = args.c5_train, args.c5_test
    t0 = time.perf_counter()

    def prog(k):
        log(tag, "training data: %d/%d problems solved, %.1f s" % (k, ntrain, time.perf_counter() - t0))

    feat, targ, _ = learned.generate(list(range(ntrain)), n=32, maxk=3, chunk=128, progress=prog)
    gen_s = time.perf_counter() - t0
    g = learned.LearnedGuess(seed=3)
    hist = g.fit(feat, targ, steps=3000, batch=16384, lr=2e-3)
    train_s = time.perf_counter() - t0 - gen_s
    log(tag, "trained on %d samples in %.1f s, loss %.2e" % (len(targ), train_s, float(hist[-10:].mean())))
    del feat, targ
    # the protocol's solver restarts its inner CG at every refinement step;
    # the last row is the default mixed solver (residual replacement, inner
    # tolerance 1e-2) from zero, for comparison
    kw = dict(solver=ta.TR_SOLVER_MIXED, maxit=30, tol_abs=0.0, tol_rel=1e-8, inner_maxit=20000, inner_tol=1e-5)
    rows = {k: dict(inner=[], seconds=[], converged=0, fallback=0) for k in ("zero_start", "learned_start",
                                                                          "learned_coarse", "zero_start_replacement")}
    pred_err, mean_err = [], []
    seeds = list(range(100000, 100000 + ntest))
    for c0 in range(0, ntest, 16):
        for P in learned.cube_problems(seeds[c0:c0 + 16], n=32, maxk=3):
            ctx = ta.HipContext(P.mesh, P.ang)
            for mode in rows:
                ctx.set_residual_replacement(mode == "zero_start_replacement")
                if mode == "zero_start_replacement":
                    d = ctx.allocVector()
                    st, _ = ctx.solve(d, use_prec=2, **dict(kw, inner_tol=1e-2))
                    gs = 0.0
                elif mode == "zero_start":
                    d = ctx.allocVector()
                    st, _ = ctx.solve(d, use_prec=2, **kw)
                    gs = 0.0
                    U, _ = ctx.moments(d)
                    t = U * (1 - P.albedo) / (4 * np.pi)
                    y = g.predict(ta.mlp_features(P.mesh, P.albedo))
                    pred_err.append(float(np.abs(y - t).mean()))
                    mean_err.append(float(np.abs(t.mean() - t).mean()))
                else:
                    ctx.synchronize()
                    t1 = time.perf_counter()
                    d = g.write_guess(ctx, P.mesh, P.albedo)
                    ctx.synchronize()
                    gs = time.perf_counter() - t1
                    st, _ = ctx.solve(d, use_prec=2 if mode == "learned_start" else 5, x0=True, **kw)
                r = rows[mode]
                r["inner"].append(st.preck)
                r["seconds"].append(gs + st.seconds)
                r["converged"] += st.status == 0
                r["fallback"] += st.fallback != 0
                ctx.deviceFree(d)
            ctx.close()
        log(tag, "held-out %d/%d: mean inner %s" % (min(c0 + 16, ntest), ntest,
                                                    {k: round(float(np.mean(v["inner"])), 1) for k, v in rows.items()}))
    base = rows["zero_start"]
    out = dict(metric="learned part: held-out mean time and inner iterations to a 1e-8 FP64 residual", unit="s",
               higher_is_better=False, n_gpus=1, data="synthetic", dtype="f32+f64",
               config=dict(workload="cube32 blobs: 32^3 cubes (196,608 tets), K=3, kappa=0.1+8 Gaussian blobs, "
                                    "albedo U[0,0.99]", train_problems=ntrain, held_out_problems=ntest),
               training=dict(generate_seconds=gen_s, train_seconds=train_s, loss=float(hist[-10:].mean())),
               prediction=dict(mean_abs_error=float(np.mean(pred_err)), mean_predictor_error=float(np.mean(mean_err))))
    for k, v in rows.items():
        out[k] = dict(mean_inner=float(np.mean(v["inner"])), mean_seconds=float(np.mean(v["seconds"])),
                      converged=v["converged"], fallback=v["fallback"],
                      inner_reduction=1 - float(np.mean(v["inner"])) / float(np.mean(base["inner"])),
                      time_reduction=1 - float(np.mean(v["seconds"])) / float(np.mean(base["seconds"])),
                      improved_problems=int(sum(a < b for a, b in zip(v["inner"], base["inner"]))))
    out["value"] = out["learned_start"]["mean_seconds"]
    print(json.dumps(out), flush=True)


def bench_c2_brief(steps, warmup):
    """The 64^3 Planck ball (K = 3): both operators and the plain FP64 and
    best mixed solves."""
    import transfer_amd as ta
    P = ta.problems.ball(64)
    ctx = ta.HipContext(P.mesh, P.ang)
    ctx.autotune(0)
    ctx.autotune(1)
    _, op64 = time_operator(ctx, P, 0, steps, warmup, 8)
    _, op32 = time_operator(ctx, P, 1, steps, warmup, 4)
    ctx.set_chebyshev(CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT)
    solves = run_solves(ctx, [("pcg_fp64_device", ta.TR_SOLVER_PCG_FP64, 0, {}),
                              ("mixed_fp32_ir_chebyshev", ta.TR_SOLVER_MIXED, 4, dict(inner_tol=1e-4))], "bench-c2")
    ctx.close()
    return dict(workload="%s: ball of 64^3-grid cubes, K=3, kappa=50exp(-r/0.2), I_p=(1-0.8r)^4" % P.name,
                nP=P.mesh.nP, operator_fp64=op64, operator_fp32_bf16=op32, time_to_1e_8=solves,
                speedup_mixed_vs_plain_fp64=solves["pcg_fp64_device"]["seconds"] /
                solves["mixed_fp32_ir_chebyshev"]["seconds"])


def c1_on_gpu():
    """The reference's CPU configuration (12x12x48 jittered slab, K = 1,
    isotropic albedo 0.5) solved on the GPU with the reference's algorithm
    (host-driven, with and without invprec) and the device PCG."""
    import transfer_amd as ta
    P = ta.problems.slab()
    ctx = ta.HipContext(P.mesh, P.ang)
    solves = run_solves(ctx, [("reference_pcg_fp64_host", ta.TR_SOLVER_REFERENCE, 0, {}),
                              ("reference_pcg_fp64_invprec_host", ta.TR_SOLVER_REFERENCE, 1,
                               dict(inner_tol=1e-8, inner_maxit=1000)),
                              ("pcg_fp64_device", ta.TR_SOLVER_PCG_FP64, 0, {})], "bench-c1")
    ctx.close()
    return dict(workload="%s: slab 12x12x48 cubes, jitter 0.15, K=1, albedo 0.5" % P.name, nP=P.mesh.nP,
                time_to_1e_8=solves)


def bench_c2(args):
    """The 64^3 Planck ball (K = 3) as the main workload: the FP64 operator
    as the step, every solver."""
    import transfer_amd as ta
    tag = "bench-c2"
    P = ta.problems.ball(args.n)
    ctx = ta.HipContext(P.mesh, P.ang)
    nP, nT, slm = P.mesh.nP, P.mesh.nT, P.ang.slm
    log(tag, "%s nP=%d nT=%d nF=%d slm=%d" % (P.name, nP, nT, P.mesh.nF, slm))
    ctx.autotune(0)
    if args.copy_calibration:
        cms = ctx.bench_copy(256 << 20, 10)
        log(tag, "copy 256 MiB: %.3f ms" % cms)
    wall, op64 = time_operator(ctx, P, 0, args.steps, args.warmup, 8)
    ctx.autotune(1)
    _, op32 = time_operator(ctx, P, 1, args.steps, args.warmup, 4)
    traffic, tsrc = measured_traffic(op64["kernel"], P.name)
    out = dict(metric=METRIC, value=op64["dof_ordinate_per_s"], unit="DOF*ordinate/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=wall / args.steps * 1e3, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="%s: ball of 64^3-grid cubes, K=3 (slm 28), kappa=50exp(-r/0.2), "
                                    "I_p=(1-0.8r)^4, FP64 operator apply" % P.name, nP=nP, nT=nT, nF=P.mesh.nF,
                           slm=slm))
    out["roofline"] = dict(bound="hbm", achieved=op64["achieved_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=op64["achieved_gbs"] / HBM_PEAK_GBS, traffic=traffic, traffic_source=tsrc,
                           kernel=op64["kernel"], kernel_ms=op64["kernel_ms"],
                           algorithmic_bytes=op64["algorithmic_bytes"])
    out["operator_fp32_bf16"] = op32
    if not args.no_solve:
        ctx.set_chebyshev(CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT)
        runs = [("reference_pcg_fp64_host", ta.TR_SOLVER_REFERENCE, 0, {}),
                ("pcg_fp64_device", ta.TR_SOLVER_PCG_FP64, 0, {}),
                ("pcg_fp64_device_jacobi", ta.TR_SOLVER_PCG_FP64, 2, {}),
                ("pcg_fp64_device_multilevel", ta.TR_SOLVER_PCG_FP64, 3, {}),
                ("pcg_fp64_device_chebyshev", ta.TR_SOLVER_PCG_FP64, 4, {}),
                ("mixed_fp32_ir_jacobi", ta.TR_SOLVER_MIXED, 2, {}),
                ("mixed_fp32_ir_multilevel", ta.TR_SOLVER_MIXED, 3, dict(inner_tol=1e-4)),
                ("mixed_fp32_ir_chebyshev", ta.TR_SOLVER_MIXED, 4, dict(inner_tol=1e-4))]
        solves = run_solves(ctx, runs, tag, args.only)
        if not args.no_learned and not args.only:
            train = [ta.problems.ball(n, maxk=3, seed=s) for n, s in ((16, 1), (20, 2), (24, 3), (28, 4))]
            g, info = train_guess(train, tag)
            lm = learned_mixed(ctx, P, g, dict(use_prec=4, inner_tol=1e-4), tag)
            lm.update(info)
            solves["learned_mixed"] = lm
        out["time_to_1e-8"] = solves
        if not args.only:
            speedups(out, solves)
    if not args.no_cpu and not args.only:
        out["cpu_baseline"] = cpu_baseline(P, ta.problems.ball(32), ta.problems.slab())
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the ball64 and slab lines of the default run")
    ap.add_argument("--invprec", action="store_true",
                    help="also time the reference's invprec-preconditioned PCG on the main workload")
    ap.add_argument("--n", type=int, default=64, help="--c2: cubes per axis of the Planck ball")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also time a plain 256 MiB device copy (counter calibration)")
    ap.add_argument("--c2", action="store_true", help="run the 64^3 Planck ball (K = 3) as the main workload")
    ap.add_argument("--c3", action="store_true", help="the default workload (accepted for older command lines)")
    ap.add_argument("--only", default="", help="run only the named time_to_1e-8 solve (for short profiles)")
    ap.add_argument("--sphere", type=int, default=0,
                    help="run the curved sphere workload with this many cubes per axis instead (48: the measured one)")
    ap.add_argument("--batch", type=int, default=0,
                    help="run the batching workload with this many problems instead (256 = the full one)")
    ap.add_argument("--c5", action="store_true", help="run the learned part's protocol instead (cube32 blobs)")
    ap.add_argument("--c5-train", type=int, default=2000, help="--c5: solver-generated training problems")
    ap.add_argument("--c5-test", type=int, default=200, help="--c5: held-out problems")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.batch:
        bench_batch(args.batch, args.steps, args.warmup, not args.no_solve, args.copy_calibration)
    elif args.sphere:
        bench_sphere(args.sphere, min(args.steps, 50), min(args.warmup, 5))
    elif args.c2:
        bench_c2(args)
    elif args.c5:
        bench_c5(args)
    else:
        bench_c3(args)


if __name__ == "__main__":
    main()
